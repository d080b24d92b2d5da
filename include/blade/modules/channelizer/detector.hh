// Channelizer -> [Polarizer XY -> LR] -> Detector as one module over the
// fused kernel of blade_channelizer_detector: the spectra never reach
// memory.  No reference counterpart as a module: the Mode-H bundle
// (reference include/blade/bundles/generic/mode_h.hh:55-106) chains the
// three modules, whose intermediate buffers are private to it, and uses this
// instead.  Validation and stepped integration follow the Channelizer
// (src/modules/channelizer/base.cc:17-42) and the Detector
// (src/modules/detector/base.cc:17-51) of this library.
#ifndef BLADE_MODULES_CHANNELIZER_DETECTOR_HH
#define BLADE_MODULES_CHANNELIZER_DETECTOR_HH

#include <string>

#include "blade/module.hh"

namespace Blade::Modules {

template <typename IT, typename OT>
class BLADE_API ChannelizerDetector : public Module {
 public:
    struct Config {
        U64 rate;
        BOOL convertToCircular = false;
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
    std::string name() const override { return "ChannelizerDetector"; }

    explicit ChannelizerDetector(const Config& config, const Input& input, const Stream& stream = {});
    Result process(const U64& currentStepCount, const Stream& stream = {}) override;

    // Rates the fused kernel runs: C <= 8192 made of the factors 2, 3, 5, 7.
    static bool Supports(U64 rate);

 private:
    const Config config;
    const Input input;
    Output output;
    U64 steps;

    U64 spectra() const { return getInputBuffer().shape().numberOfTimeSamples() / config.rate; }
    const ArrayShape getOutputBufferShape() const {
        return ArrayShape({
            getInputBuffer().shape().numberOfAspects(),
            getInputBuffer().shape().numberOfFrequencyChannels() * config.rate,
            (steps > 1) ? U64(1) : spectra() / config.integrationSize,
            config.numberOfOutputPolarizations,
        });
    }
};

}  // namespace Blade::Modules

#endif
