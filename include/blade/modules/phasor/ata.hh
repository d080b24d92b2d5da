// ATA phasor module (reference include/blade/modules/phasor/ata.hh:18-43,
// src/modules/phasor/ata.cc:53-213): per block, geometric delays of every
// (beam, antenna) relative to the boresight are computed on the host in
// FP64, uploaded, and the phasors exp(-j2pi tau f) * cal are generated on the
// GPU with an FP64 phase argument.
#ifndef BLADE_MODULES_PHASOR_ATA_HH
#define BLADE_MODULES_PHASOR_ATA_HH

#include <vector>

#include "blade/modules/phasor/generic.hh"

namespace Blade::Modules::Phasor {

template <typename OT>
class BLADE_API ATA : public Generic<OT> {
 public:
    explicit ATA(const typename Generic<OT>::Config& config,
                 const typename Generic<OT>::Input& input, const Stream& stream = {});
    ~ATA();
    Result process(const U64& currentStepCount, const Stream& stream = {}) final;

 protected:
    const PhasorShape getOutputPhasorsShape() const {
        return PhasorShape({this->config.beamCoordinates.size(), this->config.numberOfAntennas,
                            this->config.numberOfFrequencyChannels, 1,
                            this->config.numberOfPolarizations});
    }
    const DelayShape getOutputDelaysShape() const {
        return DelayShape({this->config.beamCoordinates.size(), this->config.numberOfAntennas});
    }
    const ArrayShape getConfigCalibrationShape() const {
        return ArrayShape({this->config.numberOfAntennas,
                           this->config.numberOfFrequencyChannels * this->numberOfFrequencySteps, 1,
                           this->config.numberOfPolarizations});
    }

 private:
    std::vector<XYZ> antennasXyz;
    hipEvent_t uploaded = nullptr;   // the last upload of the host delays has finished
    DelayTensor<Device::CUDA, F64> deviceDelays;
    ArrayTensor<Device::CUDA, CF64> deviceCalibrations;
};

}  // namespace Blade::Modules::Phasor

#endif
