// Module base class (reference include/blade/module.hh:20-115).  The
// modules of this library launch kernels compiled ahead of time for gfx950
// through the C-ABI of blade_amd.h and use the default constructor.  A
// module written against the reference's base compiles its kernels at run
// time instead: it passes a Program (hipRTC; the reference passes a
// jitify2::PreprocessedProgram) to the constructor, then declares kernels
// with createKernel() and launches them with runKernel(), with the
// reference's signatures.
#ifndef BLADE_MODULE_HH
#define BLADE_MODULE_HH

#include <hip/hip_runtime_api.h>

#include <string>
#include <type_traits>
#include <unordered_map>

#include "blade/macros.hh"
#include "blade/memory/base.hh"
#include "blade/program.hh"

namespace Blade {

class BLADE_API Module {
 public:
    Module() = default;
    explicit Module(const Program& program) : program(program) {}
    virtual ~Module() = default;

    virtual Taint getTaint() const { return Taint::NONE; }
    virtual U64 getComputeRatio() const { return 1; }
    virtual std::string name() const = 0;
    virtual Result process(const U64& currentStepNumber, const Stream& stream = {}) {
        return Result::SUCCESS;
    }
    // False for modules whose process() does host work per call (a replayed
    // hipGraph would repeat the device work of the captured call only).
    virtual bool graphCapturable() const { return true; }

 protected:
    // Declare kernel `name`: the instantiation key<templateArguments...> of
    // the program's __global__ template `key` (type names such as
    // TypeInfo<T>::name, or integral values), launched on gridSize x
    // blockSize.  Compiles it now (reference module.hh:40-63).
    template <typename... Args>
    Result createKernel(const std::string& name, const std::string& key, const dim3& gridSize,
                        const dim3& blockSize, const Args... templateArguments) {
        if (blockSize.x * blockSize.y * blockSize.z > 1024) {
            BL_FATAL("The block size ({}, {}, {}) is larger than hardware limit (1024).", blockSize.x,
                     blockSize.y, blockSize.z);
            return Result::ERROR;
        }
        if ((blockSize.x % 64) != 0) {
            BL_WARN("Best performance is achieved when the block size ({}, {}, {}) is a multiple of 64.",
                    blockSize.x, blockSize.y, blockSize.z);
        }
        if (program.empty()) {
            BL_FATAL("Module '{}' declares kernel '{}' without a program.", this->name(), name);
            return Result::ERROR;
        }
        std::string expression = key;
        if constexpr (sizeof...(Args) > 0) {
            std::string list;
            ((list += (list.empty() ? "" : ", ") + TemplateArgument(templateArguments)), ...);
            expression += "<" + list + ">";
        }
        hipFunction_t fn = nullptr;
        BL_CHECK(program.getKernel(expression, &fn));
        kernels[name] = Kernel{gridSize, blockSize, expression, fn};
        return Result::SUCCESS;
    }

    // Launch kernel `name` on `stream` with the arguments passed by value
    // (tensors arrive as device views of the same memory; reference
    // module.hh:65-81).
    template <typename... Args>
    Result runKernel(const std::string& name, const Stream& stream, Args... kernelArguments) {
        const auto it = kernels.find(name);
        if (it == kernels.end()) {
            BL_FATAL("Module '{}' has no kernel '{}'.", this->name(), name);
            return Result::ERROR;
        }
        const Kernel& k = it->second;
        void* params[sizeof...(Args) + 1] = {static_cast<void*>(&kernelArguments)..., nullptr};
        const hipError_t e = hipModuleLaunchKernel(k.fn, k.gridSize.x, k.gridSize.y, k.gridSize.z,
                                                   k.blockSize.x, k.blockSize.y, k.blockSize.z, 0,
                                                   stream.handle, params, nullptr);
        if (e != hipSuccess) {
            BL_FATAL("Module failed to execute: {}", hipGetErrorString(e));
            return Result::CUDA_ERROR;
        }
        return Result::SUCCESS;
    }

    static dim3 PadGridSize(const dim3& gridSize, const dim3& blockSize) {
        return dim3((gridSize.x + (blockSize.x - 1)) / blockSize.x,
                    (gridSize.y + (blockSize.y - 1)) / blockSize.y,
                    (gridSize.z + (blockSize.z - 1)) / blockSize.z);
    }

    // Alias `dst` to `src` (in-place modules), optionally with a new shape.
    template <Device D, typename T, typename S>
    static Result Link(Vector<D, T, S>& dst, const Vector<D, T, S>& src) {
        dst = src;
        return Result::SUCCESS;
    }

    template <Device D, typename T, typename S>
    static Result Link(Vector<D, T, S>& dst, const Vector<D, T, S>& src, const S& dstShape) {
        dst = src;
        return dst.reshape(dstShape);
    }

 private:
    struct Kernel {
        dim3 gridSize;
        dim3 blockSize;
        std::string key;
        hipFunction_t fn;
    };

    Program program;
    std::unordered_map<std::string, Kernel> kernels;

    template <typename T>
    static std::string TemplateArgument(const T& v) {
        if constexpr (std::is_same_v<T, bool>) {
            return v ? "true" : "false";
        } else if constexpr (std::is_integral_v<T>) {
            return std::to_string(v);
        } else {
            return std::string(v);
        }
    }
};

}  // namespace Blade

#endif
