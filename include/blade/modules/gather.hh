// Gather module (reference include/blade/modules/gather.hh:17-94): tiles
// `multiplier` successive input blocks along `axis` into one output; CHRONOUS
// with compute ratio `multiplier`.
#ifndef BLADE_MODULES_GATHER_HH
#define BLADE_MODULES_GATHER_HH

#include <string>

#include "blade/module.hh"

namespace Blade::Modules {

template <typename IT, typename OT>
class BLADE_API Gather : public Module {
 public:
    struct Config {
        U64 axis = 2;
        U64 multiplier = 1;
        U64 copySizeThreshold = 512;
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

    Taint getTaint() const override { return Taint::CONSUMER | Taint::PRODUCER | Taint::CHRONOUS; }
    U64 getComputeRatio() const override { return computeRatio; }
    std::string name() const override { return "Gather"; }

    explicit Gather(const Config& config, const Input& input, const Stream& stream = {});
    Result process(const U64& currentStepCount, const Stream& stream = {}) override;

 private:
    const Config config;
    const Input input;
    Output output;
    U64 computeRatio;
    U64 widthSize, widthByteSize, heightSize;

    const ArrayShape getOutputBufferShape() const {
        ArrayShape s = getInputBuffer().shape();
        s[config.axis] *= config.multiplier;
        return ArrayShape(s);
    }
};

}  // namespace Blade::Modules

#endif
