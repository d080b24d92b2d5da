// Run-time compilation of module kernels with hipRTC (the reference compiles
// its kernels with NVRTC through jitify2: include/blade/module.hh:22-83,
// src/kernels/base.cc).  Compiled instantiations are cached per process,
// device and name expression, and their code objects stay loaded until the
// process ends.
#include "blade/base.hh"
#include "blade/program.hh"

#include <hip/hiprtc.h>

#include <cstddef>
#include <fstream>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>

namespace Blade {

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "PROGRAM"

namespace {

// "blade/memory/base.hh" as seen by a kernel compiled at run time: the
// device side of include/blade/memory/{types,shape,custom,vector}.hh.  The
// tensor views keep the host Vector's layout (shape, data, reference count,
// unified flag), checked below, so a host tensor passed by value to a kernel
// arrives as a view of the same memory.  ops::complex is the interleaved
// {re, im} complex of the reference's device code
// (include/blade/memory/device/ops.hh:20-90).
constexpr const char* kDeviceBase = R"BLADE(
#pragma once
namespace Blade {

typedef _Float16 F16;
typedef float F32;
typedef double F64;
typedef signed char I8;
typedef short I16;
typedef int I32;
typedef long long I64;
typedef unsigned char U8;
typedef unsigned short U16;
typedef unsigned int U32;
typedef unsigned long long U64;
typedef bool BOOL;

namespace ops {

template <typename T>
class alignas(2 * sizeof(T)) complex {
 public:
    __host__ __device__ constexpr complex() : _real(0), _imag(0) {}
    __host__ __device__ constexpr complex(T r) : _real(r), _imag(0) {}
    __host__ __device__ constexpr complex(T r, T i) : _real(r), _imag(i) {}
    __host__ __device__ constexpr T real() const { return _real; }
    __host__ __device__ constexpr T imag() const { return _imag; }
    __host__ __device__ complex operator+(const complex& o) const { return complex(_real + o._real, _imag + o._imag); }
    __host__ __device__ complex operator-(const complex& o) const { return complex(_real - o._real, _imag - o._imag); }
    __host__ __device__ complex operator*(const complex& o) const {
        return complex(_real * o._real - _imag * o._imag, _real * o._imag + _imag * o._real);
    }
    __host__ __device__ complex operator/(const complex& o) const {
        const T d = o._real * o._real + o._imag * o._imag;
        return complex((_real * o._real + _imag * o._imag) / d, (_imag * o._real - _real * o._imag) / d);
    }
    __host__ __device__ bool operator==(const complex& o) const { return _real == o._real && _imag == o._imag; }
    __host__ __device__ bool operator!=(const complex& o) const { return !(*this == o); }

 private:
    T _real;
    T _imag;
};

}  // namespace ops

typedef ops::complex<F16> CF16;
typedef ops::complex<F32> CF32;
typedef ops::complex<F64> CF64;
typedef ops::complex<I8> CI8;
typedef ops::complex<I32> CI32;

enum class Device : unsigned char { NONE = 0, CPU = 1 << 0, CUDA = 1 << 1 };

template <U64 D>
struct Shape {
    U64 _shape[D];
    U64 _strides[D];
    U64 _size;

    __host__ __device__ U64 size() const { return _size; }
    __host__ __device__ static constexpr U64 dimensions() { return D; }
    __host__ __device__ const U64& operator[](U64 i) const { return _shape[i]; }
    __host__ __device__ U64 stride(U64 i) const { return _strides[i]; }
    __host__ __device__ U64 shapeToOffset(const U64 (&c)[D]) const {
        U64 index = 0;
        for (U64 i = 0; i < D; i++) index += c[i] * _strides[i];
        return index;
    }
};

struct ArrayShape : Shape<4> {
    __host__ __device__ U64 numberOfAspects() const { return _shape[0]; }
    __host__ __device__ U64 numberOfFrequencyChannels() const { return _shape[1]; }
    __host__ __device__ U64 numberOfTimeSamples() const { return _shape[2]; }
    __host__ __device__ U64 numberOfPolarizations() const { return _shape[3]; }
};

struct PhasorShape : Shape<5> {
    __host__ __device__ U64 numberOfBeams() const { return _shape[0]; }
    __host__ __device__ U64 numberOfAntennas() const { return _shape[1]; }
    __host__ __device__ U64 numberOfFrequencyChannels() const { return _shape[2]; }
    __host__ __device__ U64 numberOfTimeSamples() const { return _shape[3]; }
    __host__ __device__ U64 numberOfPolarizations() const { return _shape[4]; }
};

struct DelayShape : Shape<2> {
    __host__ __device__ U64 numberOfBeams() const { return _shape[0]; }
    __host__ __device__ U64 numberOfAntennas() const { return _shape[1]; }
};

struct VectorShape : Shape<1> {};

template <Device DeviceId, typename DataType, typename S>
struct Vector {
    S _shape;
    DataType* _data;
    void* _refs;
    bool _unified;

    __host__ __device__ DataType* data() const { return _data; }
    __host__ __device__ U64 size() const { return _shape.size(); }
    __host__ __device__ U64 size_bytes() const { return _shape.size() * sizeof(DataType); }
    __host__ __device__ const S& shape() const { return _shape; }
    __host__ __device__ DataType& operator[](U64 i) const { return _data[i]; }
};

template <Device D, typename T> using ArrayTensor = Vector<D, T, ArrayShape>;
template <Device D, typename T> using PhasorTensor = Vector<D, T, PhasorShape>;
template <Device D, typename T> using DelayTensor = Vector<D, T, DelayShape>;
template <Device D, typename T> using Tensor = Vector<D, T, VectorShape>;

}  // namespace Blade
)BLADE";

// The device view above assumes this host layout.
using HostView = ArrayTensor<Device::CUDA, CF32>;
static_assert(sizeof(ArrayShape) == 9 * sizeof(U64), "ArrayShape layout");
static_assert(sizeof(HostView) == 9 * sizeof(U64) + 3 * sizeof(void*), "tensor layout");

std::mutex g_cache_mutex;
std::map<std::string, hipFunction_t> g_kernels;

std::string hiprtc_error(const std::string& what, hiprtcResult r) {
    return what + ": " + hiprtcGetErrorString(r);
}

}  // namespace

Program::Program(std::string name, std::string source, std::vector<std::string> options)
    : _name(std::move(name)), _source(std::move(source)), _options(std::move(options)) {}

Program Program::FromFile(const std::string& path, std::vector<std::string> options) {
    std::ifstream f(path);
    if (!f) {
        BL_FATAL("Cannot read kernel source '{}'.", path);
        throw Result::ERROR;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const size_t slash = path.find_last_of('/');
    return Program(slash == std::string::npos ? path : path.substr(slash + 1), ss.str(), std::move(options));
}

Result Program::compile(const std::string& arch, const std::string& nameExpression, std::string* loweredName,
                        std::vector<char>* code, std::string* log) const {
    const char* headers[] = {kDeviceBase};
    const char* includeNames[] = {"blade/memory/base.hh"};
    hiprtcProgram prog = nullptr;
    hiprtcResult r = hiprtcCreateProgram(&prog, _source.c_str(), _name.c_str(), 1, headers, includeNames);
    if (r != HIPRTC_SUCCESS) {
        if (log) *log = hiprtc_error("hiprtcCreateProgram", r);
        return Result::ERROR;
    }
    Result result = Result::SUCCESS;
    std::vector<std::string> opts = {"--offload-arch=" + arch, "-O3", "-std=c++17",
                                     // no packed FP32 instructions in any HIP code of this library
                                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    opts.insert(opts.end(), _options.begin(), _options.end());
    std::vector<const char*> copts;
    for (const auto& o : opts) copts.push_back(o.c_str());
    r = hiprtcAddNameExpression(prog, nameExpression.c_str());
    if (r == HIPRTC_SUCCESS) r = hiprtcCompileProgram(prog, static_cast<int>(copts.size()), copts.data());
    if (log) {
        size_t n = 0;
        if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
            log->assign(n, '\0');
            hiprtcGetProgramLog(prog, log->data());
            log->resize(n - 1);
        } else {
            log->clear();
        }
        if (r != HIPRTC_SUCCESS) *log = hiprtc_error("compiling " + _name, r) + "\n" + *log;
    }
    if (r != HIPRTC_SUCCESS) {
        result = Result::ERROR;
    } else {
        const char* lowered = nullptr;
        size_t size = 0;
        if (hiprtcGetLoweredName(prog, nameExpression.c_str(), &lowered) != HIPRTC_SUCCESS || !lowered ||
            hiprtcGetCodeSize(prog, &size) != HIPRTC_SUCCESS || size == 0) {
            if (log) *log += "\nno kernel '" + nameExpression + "' in " + _name;
            result = Result::ERROR;
        } else {
            if (loweredName) *loweredName = lowered;
            if (code) {
                code->resize(size);
                if (hiprtcGetCode(prog, code->data()) != HIPRTC_SUCCESS) result = Result::ERROR;
            }
        }
    }
    hiprtcDestroyProgram(&prog);
    return result;
}

Result Program::getKernel(const std::string& nameExpression, hipFunction_t* kernel) const {
    int device = 0;
    BL_HIP_CHECK(hipGetDevice(&device));
    std::string options;
    for (const auto& o : _options) options += o + '\n';
    const std::string key = std::to_string(device) + '\n' + _name + '\n' +
                            std::to_string(std::hash<std::string>{}(_source)) + '\n' + options + nameExpression;
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    const auto it = g_kernels.find(key);
    if (it != g_kernels.end()) {
        *kernel = it->second;
        return Result::SUCCESS;
    }
    hipDeviceProp_t props;
    BL_HIP_CHECK(hipGetDeviceProperties(&props, device));
    std::string lowered, log;
    std::vector<char> code;
    if (compile(props.gcnArchName, nameExpression, &lowered, &code, &log) != Result::SUCCESS) {
        BL_FATAL("Kernel '{}' of '{}' failed to compile:\n{}", nameExpression, _name, log);
        return Result::ERROR;
    }
    hipModule_t module = nullptr;
    BL_HIP_CHECK(hipModuleLoadData(&module, code.data()));
    hipFunction_t fn = nullptr;
    BL_HIP_CHECK(hipModuleGetFunction(&fn, module, lowered.c_str()));
    g_kernels.emplace(key, fn);
    *kernel = fn;
    return Result::SUCCESS;
}

}  // namespace Blade
