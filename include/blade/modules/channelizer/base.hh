// Channelizer module (reference include/blade/modules/channelizer/base.hh:17-86).
// Any rate dividing T is accepted (the reference accepts only 1 and T,
// src/modules/channelizer/base.cc:30-35).  The output aliases the input
// buffer for every rate, as the reference links them (base.cc:45): rate T
// transforms in place, other rates go through a staging buffer of the
// module.  A Pipeline fuses a rate of 2 or 4 into a Beamformer before it.
#ifndef BLADE_MODULES_CHANNELIZER_BASE_HH
#define BLADE_MODULES_CHANNELIZER_BASE_HH

#include <string>

#include "blade/module.hh"

namespace Blade::Modules {

template <typename IT, typename OT>
class BLADE_API Channelizer : public Module {
 public:
    struct Config {
        U64 rate = 4;
        U64 blockSize = 512;
    };

    const Config& getConfig() const { return config; }

    struct Input {
        const ArrayTensor<Device::CUDA, IT>& buf;
    };

    const ArrayTensor<Device::CUDA, IT>& getInputBuffer() const { return input.buf; }

    struct Output {
        ArrayTensor<Device::CUDA, OT> buf;
    };

    const ArrayTensor<Device::CUDA, OT>& getOutputBuffer() const { return output.buf; }

    Taint getTaint() const override { return Taint::CONSUMER | Taint::MODIFIER; }
    std::string name() const override { return "Channelizer"; }

    explicit Channelizer(const Config& config, const Input& input, const Stream& stream = {});
    Result process(const U64& currentStepCount, const Stream& stream = {}) override;

    // Pipeline fusion (no reference counterpart): the producer of the input
    // buffer runs this transform in its own launch; process() does nothing.
    void fuseIntoProducer();
    bool fused() const { return isFused; }

 private:
    const Config config;
    const Input input;
    Output output;
    bool isFused;
    // Out-of-place result for 1 < rate < T, copied back over the buffer.
    ArrayTensor<Device::CUDA, OT> staging;

    const ArrayShape getOutputBufferShape() const {
        return ArrayShape({
            getInputBuffer().shape().numberOfAspects(),
            getInputBuffer().shape().numberOfFrequencyChannels() * config.rate,
            getInputBuffer().shape().numberOfTimeSamples() / config.rate,
            getInputBuffer().shape().numberOfPolarizations(),
        });
    }
};

}  // namespace Blade::Modules

#endif
