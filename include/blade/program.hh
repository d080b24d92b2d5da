// A HIP kernel source compiled at run time with hipRTC: what a module hands
// to the Module constructor, in place of the reference's
// jitify2::PreprocessedProgram (include/blade/module.hh:22).
//
// Kernel sources may include "blade/memory/base.hh": inside a compiled
// program that name is the device side of this API (the scalar and complex
// types F32, CF32, CI8 ..., ArrayShape and the other shapes, and tensor views
// ArrayTensor<Device::CUDA, T> ... with the host tensors' memory layout,
// so a host tensor passed to runKernel() arrives as the same tensor).
// Compiled code is cached per process, device and template instantiation.
#ifndef BLADE_PROGRAM_HH
#define BLADE_PROGRAM_HH

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "blade/macros.hh"

namespace Blade {

class BLADE_API Program {
 public:
    Program() = default;
    // `name` is the source's file name in compiler messages; `options` are
    // added to the compiler's (-O3 -std=c++17, the device's architecture).
    Program(std::string name, std::string source, std::vector<std::string> options = {});
    static Program FromFile(const std::string& path, std::vector<std::string> options = {});

    const std::string& name() const { return _name; }
    const std::string& source() const { return _source; }
    const std::vector<std::string>& options() const { return _options; }
    bool empty() const { return _source.empty(); }

    // Compile one instantiation (`nameExpression`, e.g. "scale<CF32, 2>")
    // for `arch` (e.g. "gfx950"); needs no device.  On failure the compiler
    // log is in `log`.
    Result compile(const std::string& arch, const std::string& nameExpression, std::string* loweredName,
                   std::vector<char>* code, std::string* log) const;

    // The kernel of `nameExpression` on the current device, compiled and
    // loaded on first use.
    Result getKernel(const std::string& nameExpression, hipFunction_t* kernel) const;

 private:
    std::string _name;
    std::string _source;
    std::vector<std::string> _options;
};

}  // namespace Blade

#endif
