// Detector module (reference include/blade/modules/detector.hh:17-80).
// When integrationSize is a multiple of the block's time samples the module
// integrates across ceil(integrationSize / T) consecutive blocks (stepped
// integration, CHRONOUS with that compute ratio); the reference integrates
// only within a block (detector.hh:11).
#ifndef BLADE_MODULES_DETECTOR_HH
#define BLADE_MODULES_DETECTOR_HH

#include <string>

#include "blade/module.hh"

namespace Blade::Modules {

template <typename IT, typename OT>
class BLADE_API Detector : public Module {
 public:
    struct Config {
        U64 integrationSize;
        U64 numberOfOutputPolarizations = 4;
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

    Taint getTaint() const override {
        return (steps > 1) ? (Taint::CONSUMER | Taint::PRODUCER | Taint::CHRONOUS)
                           : (Taint::CONSUMER | Taint::PRODUCER);
    }
    U64 getComputeRatio() const override { return steps; }
    std::string name() const override { return "Detector"; }

    explicit Detector(const Config& config, const Input& input, const Stream& stream = {});
    Result process(const U64& currentStepCount, const Stream& stream = {}) override;

 private:
    const Config config;
    const Input input;
    Output output;
    U64 steps;

    const ArrayShape getOutputBufferShape() const {
        const U64 T = getInputBuffer().shape().numberOfTimeSamples();
        return ArrayShape({
            getInputBuffer().shape().numberOfAspects(),
            getInputBuffer().shape().numberOfFrequencyChannels(),
            (steps > 1) ? U64(1) : T / config.integrationSize,
            config.numberOfOutputPolarizations,
        });
    }
};

}  // namespace Blade::Modules

#endif
