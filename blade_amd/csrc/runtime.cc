// Library identification and per-thread error messages of the C-ABI.
#include <string>

#include "blade_amd.h"

namespace bl {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace bl

extern "C" const char* blade_amd_version(void) {
    return "blade_amd 0.1.0 (gfx950)";
}

extern "C" const char* blade_amd_last_error(void) {
    return bl::g_last_error.c_str();
}
