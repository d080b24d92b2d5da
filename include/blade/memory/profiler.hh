// Allocation profiler (reference include/blade/memory/profiler.hh:9-90): counts
// bytes allocated per memory space; while capturing, tensors record their
// size but allocate nothing.
#ifndef BLADE_MEMORY_PROFILER_HH
#define BLADE_MEMORY_PROFILER_HH

#include <atomic>

#include "blade/memory/types.hh"

namespace Blade {

class BLADE_API Profiler {
 public:
    static Profiler& Get();

    static void StartCapture() { Get().capturing = true; Get().reset(); }
    static void StopCapture() { Get().capturing = false; }
    static bool IsCapturing() { return Get().capturing.load(); }

    static void RegisterCpuAllocation(U64 b) { Get().cpuAlloc += b; Get().cpuCur += b; }
    static void RegisterCudaAllocation(U64 b) { Get().devAlloc += b; Get().devCur += b; }
    static void RegisterUnifiedAllocation(U64 b) { Get().uniAlloc += b; Get().uniCur += b; }
    static void RegisterCpuDeallocation(U64 b) { Get().cpuCur -= b; }
    static void RegisterCudaDeallocation(U64 b) { Get().devCur -= b; }
    static void RegisterUnifiedDeallocation(U64 b) { Get().uniCur -= b; }

    static U64 GetCpuAllocations() { return Get().cpuAlloc; }
    static U64 GetCudaAllocations() { return Get().devAlloc; }
    static U64 GetUnifiedAllocations() { return Get().uniAlloc; }
    static U64 GetCpuCurrent() { return Get().cpuCur; }
    static U64 GetCudaCurrent() { return Get().devCur; }
    static U64 GetUnifiedCurrent() { return Get().uniCur; }

 private:
    std::atomic<bool> capturing{false};
    std::atomic<U64> cpuAlloc{0}, devAlloc{0}, uniAlloc{0};
    std::atomic<U64> cpuCur{0}, devCur{0}, uniCur{0};

    void reset() {
        cpuAlloc = devAlloc = uniAlloc = 0;
        cpuCur = devCur = uniCur = 0;
    }
};

}  // namespace Blade

#endif
