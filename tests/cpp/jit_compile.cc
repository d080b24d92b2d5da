// Run-time compilation without a device (Program::compile, hipRTC): a valid
// template instantiation compiles to a code object, a source error and an
// unknown kernel fail with the compiler's message.  Prints one line per case.
#include <cstdio>
#include <string>
#include <vector>

#include "blade/base.hh"

using namespace Blade;

int main() {
    const Program good("good.hip", R"(
#include "blade/memory/base.hh"
using namespace Blade;
template <typename T, int N>
__global__ void add(ArrayTensor<Device::CUDA, T> x) {
    const U64 i = (U64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < x.size()) x[i] = x[i] + T(N, -N);
}
)");
    const Program bad("bad.hip", "__global__ void broken( { }\n");
    std::string lowered, log;
    std::vector<char> code;
    int failures = 0;

    Result r = good.compile("gfx950", "add<CF32, 3>", &lowered, &code, &log);
    std::printf("good: %d lowered=%s bytes=%zu\n", (int)r, lowered.c_str(), code.size());
    failures += !(r == Result::SUCCESS && !lowered.empty() && code.size() > 0);

    r = good.compile("gfx950", "add<CI8, 1>", &lowered, &code, &log);
    std::printf("good ci8: %d lowered=%s\n", (int)r, lowered.c_str());
    failures += !(r == Result::SUCCESS);

    r = good.compile("gfx950", "missing<CF32>", &lowered, &code, &log);
    std::printf("missing: %d\n", (int)r);
    failures += !(r == Result::ERROR);

    r = bad.compile("gfx950", "broken", &lowered, &code, &log);
    std::printf("bad: %d log=%s\n", (int)r, log.find("bad.hip") != std::string::npos ? "names the file" : log.c_str());
    failures += !(r == Result::ERROR && log.find("bad.hip") != std::string::npos);
    return failures;
}
