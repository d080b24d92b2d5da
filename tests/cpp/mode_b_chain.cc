// The C++ API as a reference user writes it (reference
// benchmarks/modules/beamformer/base.hh, include/blade/runner.hh): a Runner
// with Beamformer::ATA -> Channelizer(rate 4) -> Polarizer(XY -> LR)
// connected through designated-initializer configs, host tensors copied in,
// one compute step, the beams copied out.
//
//   mode_b_chain <x.ci8> <w.cf32> <y.cf32> A F T B
//
// x: int8 [A, F, T, 2, 2]; w: complex64 [B, A, F, 1, 2]; y: complex64
// [B, 4F, T/4, 2].  Exit status 0 on success.
#include <cstdio>
#include <cstdlib>
#include <memory>

#include "blade/base.hh"
#include "blade/modules/base.hh"

using namespace Blade;

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
    if (argc != 8) {
        std::fprintf(stderr, "usage: %s x.ci8 w.cf32 y.cf32 A F T B\n", argv[0]);
        return 2;
    }
    const U64 A = std::strtoull(argv[4], nullptr, 10), F = std::strtoull(argv[5], nullptr, 10);
    const U64 T = std::strtoull(argv[6], nullptr, 10), B = std::strtoull(argv[7], nullptr, 10);
    try {
        Runner runner;
        ArrayTensor<Device::CPU, CI8> hx({A, F, T, 2});
        PhasorTensor<Device::CPU, CF32> hw({B, A, F, 1, 2});
        ArrayTensor<Device::CUDA, CI8> x({A, F, T, 2});
        PhasorTensor<Device::CUDA, CF32> w({B, A, F, 1, 2});
        if (!read_file(argv[1], hx.data(), hx.size_bytes()) || !read_file(argv[2], hw.data(), hw.size_bytes())) {
            std::fprintf(stderr, "cannot read the inputs\n");
            return 1;
        }

        std::shared_ptr<Modules::Beamformer::ATA<CI8, CF32>> beamformer;
        std::shared_ptr<Modules::Channelizer<CF32, CF32>> channelizer;
        std::shared_ptr<Modules::Polarizer<CF32, CF32>> polarizer;
        runner.connect(beamformer, {.enableIncoherentBeam = false, .enableIncoherentBeamSqrt = false,
                                    .blockSize = 512},
                       {.buf = x, .phasors = w});
        runner.connect(channelizer, {.rate = 4, .blockSize = 512}, {.buf = beamformer->getOutputBuffer()});
        runner.connect(polarizer, {.inputPolarization = POL::XY, .outputPolarization = POL::LR, .blockSize = 512},
                       {.buf = channelizer->getOutputBuffer()});

        if (Copy(x, hx, runner.stream(0)) != Result::SUCCESS) return 1;
        if (Copy(w, hw, runner.stream(0)) != Result::SUCCESS) return 1;
        if (runner.compute(0) != Result::SUCCESS) return 1;
        ArrayTensor<Device::CPU, CF32> hy(polarizer->getOutputBuffer().shape());
        if (Copy(hy, polarizer->getOutputBuffer(), runner.stream(0)) != Result::SUCCESS) return 1;
        if (runner.synchronize(0) != Result::SUCCESS) return 1;
        if (!write_file(argv[3], hy.data(), hy.size_bytes())) return 1;
    } catch (const Result& r) {
        std::fprintf(stderr, "blade error %d\n", static_cast<int>(r));
        return 1;
    }
    return 0;
}
