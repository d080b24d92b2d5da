// Permutation module (reference include/blade/modules/permutation.hh:16-77):
// output = transpose(input, axes = indexes).
#ifndef BLADE_MODULES_PERMUTATION_HH
#define BLADE_MODULES_PERMUTATION_HH

#include <string>

#include "blade/module.hh"

namespace Blade::Modules {

template <typename IT, typename OT>
class BLADE_API Permutation : public Module {
 public:
    struct Config {
        ArrayShape indexes = ArrayShape({0, 1, 2, 3});
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

    Taint getTaint() const override { return Taint::CONSUMER | Taint::PRODUCER; }
    std::string name() const override { return "Permutation"; }

    explicit Permutation(const Config& config, const Input& input, const Stream& stream = {});
    Result process(const U64& currentStepCount, const Stream& stream = {}) override;

 private:
    const Config config;
    const Input input;
    Output output;

    const ArrayShape getOutputBufferShape() const {
        const auto& s = getInputBuffer().shape();
        return ArrayShape({s[config.indexes[0]], s[config.indexes[1]], s[config.indexes[2]],
                           s[config.indexes[3]]});
    }
};

}  // namespace Blade::Modules

#endif
