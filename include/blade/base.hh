// Umbrella header of the Blade C++ API on gfx950.
#ifndef BLADE_BASE_HH
#define BLADE_BASE_HH

#include "blade_amd.h"
#include "blade/logger.hh"
#include "blade/macros.hh"
#include "blade/types.hh"
#include "blade/memory/base.hh"
#include "blade/module.hh"
#include "blade/bundle.hh"
#include "blade/pipeline.hh"
#include "blade/runner.hh"
#include "blade/modules/base.hh"
#include "blade/bundles/ata/mode_b.hh"
#include "blade/bundles/generic/mode_h.hh"

namespace Blade {

// Select the HIP device used by subsequent allocations and launches
// (reference include/blade/base.hh:12-17, SetCudaDevice).
inline Result SetCudaDevice(int device_id) {
    BL_HIP_CHECK(hipSetDevice(device_id));
    return Result::SUCCESS;
}

}  // namespace Blade

#endif
