// A module written the way the reference's modules are (a kernel source
// handed to the Module constructor, createKernel() in the constructor,
// runKernel() in process(); reference include/blade/module.hh:22-83,
// src/modules/polarizer/base.cc:12-71), compiled at run time with hipRTC and
// run in a Runner ahead of this library's own Polarizer.
//
//   jit_module <x.cf32> <y.cf32> <shape.u64> A F T
//
// x: complex64 [A, F, T, 2]; y = polarize(2 * factor * x) with factor 0.5;
// shape.u64: what a kernel reads of the input tensor's view (size, the four
// dimensions, the four strides).  Exit status 0 on success.
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "blade/base.hh"

using namespace Blade;

static const Program kScaleProgram("scale.hip", R"(
#include "blade/memory/base.hh"

using namespace Blade;

template <typename IT, typename OT, int N>
__global__ void scale(const ArrayTensor<Device::CUDA, IT> input, ArrayTensor<Device::CUDA, OT> output,
                      float factor) {
    const U64 tid = (U64)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < input.size()) {
        const IT v = input[tid];
        output[tid] = OT(v.real() * factor * N, v.imag() * factor * N);
    }
}

template <typename T>
__global__ void describe(const ArrayTensor<Device::CUDA, T> input, Tensor<Device::CUDA, U64> out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const ArrayShape& s = input.shape();
        out[0] = input.size();
        out[1] = s.numberOfAspects();
        out[2] = s.numberOfFrequencyChannels();
        out[3] = s.numberOfTimeSamples();
        out[4] = s.numberOfPolarizations();
        for (int i = 0; i < 4; i++) out[5 + i] = s.stride(i);
    }
}
)");

template <typename IT, typename OT>
class Scale : public Module {
 public:
    struct Config {
        F32 factor;
        U64 blockSize = 256;
    };
    struct Input {
        const ArrayTensor<Device::CUDA, IT>& buf;
    };
    struct Output {
        ArrayTensor<Device::CUDA, OT> buf;
        Tensor<Device::CUDA, U64> description;
    };

    const ArrayTensor<Device::CUDA, OT>& getOutputBuffer() const { return output.buf; }
    const Tensor<Device::CUDA, U64>& getDescription() const { return output.description; }
    std::string name() const override { return "Scale"; }

    Scale(const Config& config, const Input& input, const Stream& stream = {})
        : Module(kScaleProgram), config(config), input(input) {
        output.buf = ArrayTensor<Device::CUDA, OT>(input.buf.shape());
        output.description = Tensor<Device::CUDA, U64>({9});
        BL_CHECK_THROW(createKernel("main", "scale", PadGridSize(input.buf.size(), config.blockSize),
                                    config.blockSize, TypeInfo<IT>::name, TypeInfo<OT>::name, 2));
        BL_CHECK_THROW(createKernel("describe", "describe", 1, 64, TypeInfo<IT>::name));
    }

    Result process(const U64&, const Stream& stream) override {
        BL_CHECK(runKernel("main", stream, input.buf, output.buf, config.factor));
        return runKernel("describe", stream, input.buf, output.description);
    }

 private:
    const Config config;
    const Input input;
    Output output;
};

static bool read_file(const char* path, void* dst, size_t bytes) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    const size_t n = std::fread(dst, 1, bytes, f);
    std::fclose(f);
    return n == bytes;
}

static bool write_file(const char* path, const void* src, size_t bytes) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const size_t n = std::fwrite(src, 1, bytes, f);
    std::fclose(f);
    return n == bytes;
}

int main(int argc, char** argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s x.cf32 y.cf32 shape.u64 A F T\n", argv[0]);
        return 2;
    }
    const U64 A = std::strtoull(argv[4], nullptr, 10), F = std::strtoull(argv[5], nullptr, 10);
    const U64 T = std::strtoull(argv[6], nullptr, 10);
    try {
        Runner runner;
        ArrayTensor<Device::CPU, CF32> hx({A, F, T, 2});
        ArrayTensor<Device::CUDA, CF32> x({A, F, T, 2});
        if (!read_file(argv[1], hx.data(), hx.size_bytes())) {
            std::fprintf(stderr, "cannot read the input\n");
            return 1;
        }
        std::shared_ptr<Scale<CF32, CF32>> scale;
        std::shared_ptr<Modules::Polarizer<CF32, CF32>> polarizer;
        runner.connect(scale, {.factor = 0.5f}, {.buf = x});
        runner.connect(polarizer, {.inputPolarization = POL::XY, .outputPolarization = POL::LR},
                       {.buf = scale->getOutputBuffer()});
        if (Copy(x, hx, runner.stream(0)) != Result::SUCCESS) return 1;
        if (runner.compute(0) != Result::SUCCESS) return 1;
        ArrayTensor<Device::CPU, CF32> hy(polarizer->getOutputBuffer().shape());
        Tensor<Device::CPU, U64> hd({9});
        if (Copy(hy, polarizer->getOutputBuffer(), runner.stream(0)) != Result::SUCCESS) return 1;
        if (Copy(hd, scale->getDescription(), runner.stream(0)) != Result::SUCCESS) return 1;
        if (runner.synchronize(0) != Result::SUCCESS) return 1;
        if (!write_file(argv[2], hy.data(), hy.size_bytes())) return 1;
        if (!write_file(argv[3], hd.data(), hd.size_bytes())) return 1;
    } catch (const Result& r) {
        std::fprintf(stderr, "blade error %d\n", static_cast<int>(r));
        return 1;
    }
    return 0;
}
