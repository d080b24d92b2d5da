// Bundle: an ordered group of modules built together (reference include/blade/bundle.hh:14-36).
#ifndef BLADE_BUNDLE_HH
#define BLADE_BUNDLE_HH

#include <memory>
#include <vector>

#include "blade/module.hh"

namespace Blade {

class BLADE_API Bundle {
 public:
    explicit Bundle(const Stream& stream) : stream(stream) {}
    virtual ~Bundle() = default;

    std::vector<std::shared_ptr<Module>>& getModules() { return modules; }

 protected:
    template <typename Block>
    void connect(std::shared_ptr<Block>& module, const typename Block::Config& config,
                 const typename Block::Input& input) {
        module = std::make_shared<Block>(config, input, stream);
        modules.push_back(module);
    }

 private:
    const Stream stream;
    std::vector<std::shared_ptr<Module>> modules;
};

}  // namespace Blade

#endif
