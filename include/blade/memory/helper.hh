// Double buffer used by the Runner (reference include/blade/memory/helper.hh:13-49):
// two tensors, one of which is "fixed" (the one modules are connected to);
// operator[] re-points the fixed handle to a slot.
#ifndef BLADE_MEMORY_HELPER_HH
#define BLADE_MEMORY_HELPER_HH

#include <vector>

#include "blade/memory/vector.hh"

namespace Blade {

template <typename T>
class Duet {
 public:
    template <typename... Args>
    explicit Duet(Args&&... args) {
        swapchain.emplace_back(std::forward<Args>(args)...);
        swapchain.emplace_back(std::forward<Args>(args)...);
        _fixed = swapchain[0];
    }

    void set(U64 index) { _fixed = swapchain.at(index); }
    operator T&() { return _fixed; }
    T& operator[](U64 index) {
        _fixed = swapchain.at(index);
        return _fixed;
    }
    const T& at(U64 index) const { return swapchain.at(index); }
    T& fixed() { return _fixed; }

 private:
    T _fixed;
    std::vector<T> swapchain;
};

}  // namespace Blade

#endif
