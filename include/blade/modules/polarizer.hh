// Polarizer module XY -> LR (reference include/blade/modules/polarizer.hh:14-70).
#ifndef BLADE_MODULES_POLARIZER_HH
#define BLADE_MODULES_POLARIZER_HH

#include <string>

#include "blade/module.hh"
#include "blade/types.hh"

namespace Blade::Modules {

template <typename IT, typename OT>
class BLADE_API Polarizer : public Module {
 public:
    struct Config {
        POL inputPolarization = POL::XY;
        POL outputPolarization = POL::LR;
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

    Taint getTaint() const override { return Taint::MODIFIER; }
    std::string name() const override { return "Polarizer"; }

    explicit Polarizer(const Config& config, const Input& input, const Stream& stream = {});
    Result process(const U64& currentStepCount, const Stream& stream = {}) override;

    // Pipeline fusion (no reference counterpart): the producer of the input
    // buffer polarizes in its own launch; process() does nothing.
    void fuseIntoProducer() { isFused = true; }
    bool fused() const { return isFused; }

 private:
    const Config config;
    const Input input;
    Output output;
    bool isFused = false;
};

}  // namespace Blade::Modules

#endif
