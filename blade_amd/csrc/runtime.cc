// Library identification, per-thread error messages of the C-ABI and the
// library-owned launch workspaces.
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "common.hh"

namespace bl {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

static std::mutex g_ws_mutex;
static std::map<std::tuple<int, hipStream_t, int>, std::pair<unsigned char*, size_t>> g_ws;

int library_workspace(int user, size_t bytes, hipStream_t s, unsigned char** out) {
    int dev = 0;
    hipError_t derr = hipGetDevice(&dev);
    if (derr != hipSuccess) {
        set_error(std::string("workspace: hipGetDevice: ") + hipGetErrorString(derr));
        return BLADE_HIP_ERROR;
    }
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    auto& e = g_ws[std::make_tuple(dev, s, user)];
    if (e.second < bytes) {
        if (e.first) {
            // The old buffer may still be read by work queued on this stream.
            if (hipStreamSynchronize(s) != hipSuccess || hipFree(e.first) != hipSuccess) {
                set_error("workspace: could not release the old buffer");
                return BLADE_HIP_ERROR;
            }
            e = {nullptr, 0};
        }
        void* p = nullptr;
        hipError_t err = hipMalloc(&p, bytes);
        if (err != hipSuccess) {
            set_error(std::string("workspace allocation: ") + hipGetErrorString(err));
            return BLADE_HIP_ERROR;
        }
        e = {static_cast<unsigned char*>(p), bytes};
    }
    *out = e.first;
    return BLADE_SUCCESS;
}

}  // namespace bl

extern "C" const char* blade_amd_version(void) {
    return "blade_amd 0.1.0 (gfx950)";
}

extern "C" const char* blade_amd_last_error(void) {
    return bl::g_last_error.c_str();
}
