// Module base class (reference include/blade/module.hh:20-115).  Modules
// here launch pre-compiled gfx950 kernels through the C-ABI of blade_amd.h
// instead of JIT-compiling templates, so the base keeps only the interface
// (taint, compute ratio, name, process) and the Link helpers.
#ifndef BLADE_MODULE_HH
#define BLADE_MODULE_HH

#include <string>

#include "blade/macros.hh"
#include "blade/memory/base.hh"

namespace Blade {

class BLADE_API Module {
 public:
    Module() = default;
    virtual ~Module() = default;

    virtual Taint getTaint() const { return Taint::NONE; }
    virtual U64 getComputeRatio() const { return 1; }
    virtual std::string name() const = 0;
    virtual Result process(const U64& currentStepNumber, const Stream& stream = {}) {
        return Result::SUCCESS;
    }

 protected:
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
};

}  // namespace Blade

#endif
