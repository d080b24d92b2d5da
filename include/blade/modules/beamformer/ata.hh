// ATA beamformer module (reference include/blade/modules/beamformer/ata.hh:14-24,
// src/modules/beamformer/ata.cc:7-72).  Instantiated for CF32 input (the
// reference's) and for CI8 input, whose widening is fused into the kernel.
#ifndef BLADE_MODULES_BEAMFORMER_ATA_HH
#define BLADE_MODULES_BEAMFORMER_ATA_HH

#include "blade/modules/beamformer/generic.hh"

namespace Blade::Modules::Beamformer {

template <typename IT, typename OT>
class BLADE_API ATA : public Generic<IT, OT> {
 public:
    explicit ATA(const typename Generic<IT, OT>::Config& config,
                 const typename Generic<IT, OT>::Input& input, const Stream& stream = {});

 protected:
    const ArrayShape getOutputBufferShape() const override {
        return ArrayShape({
            this->getInputPhasors().shape().numberOfBeams() +
                U64(this->config.enableIncoherentBeam ? 1 : 0),
            this->getInputBuffer().shape().numberOfFrequencyChannels(),
            this->getInputBuffer().shape().numberOfTimeSamples(),
            this->getInputBuffer().shape().numberOfPolarizations(),
        });
    }
};

}  // namespace Blade::Modules::Beamformer

#endif
