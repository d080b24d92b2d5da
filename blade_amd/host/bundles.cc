// Explicit instantiations of the bundles exported by the library
// (reference python/blade/ext_modeb.cc:111-114 and ext_modeh.cc).
#include "blade/base.hh"

namespace Blade::Bundles {

template class BLADE_API ATA::ModeB<CI8, CF32>;
template class BLADE_API ATA::ModeB<CI8, F32>;
template class BLADE_API ATA::ModeB<CF32, CF32>;
template class BLADE_API ATA::ModeB<CF32, F32>;

template class BLADE_API Generic::ModeH<CF32, F32>;
template class BLADE_API Generic::ModeH<CI8, F32>;

}  // namespace Blade::Bundles
