// Ownership semantics of the reference-counted tensors (the cases of the
// reference's tests/memory/vector/reference.cc:15-208): empty, owning and
// non-owning ("zombie") construction, copy and move construction, copy and
// move assignment, and tensors held in aggregates built with designated
// initializers.  Also: self-assignment, and that the last owner frees.
//
//   vector_semantics            real allocations (needs a GPU)
//   vector_semantics --capture  allocation profiler capture: sizes are
//                               recorded and nothing is allocated, so the
//                               program runs without a GPU; owned data()
//                               is then null.
// Prints one line per case and "vector semantics: N failures"; exit status 0
// when every check holds.
#include <cstdio>
#include <cstring>
#include <utility>

#include "blade/memory/base.hh"

using namespace Blade;

static int g_failures = 0;
static bool g_capture = false;

#define CHECK(cond)                                                           \
    do {                                                                      \
        if (!(cond)) {                                                        \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
            g_failures++;                                                     \
        }                                                                     \
    } while (0)

// An owning tensor has memory, except under capture, where it has none.
template <typename V>
static bool owns_memory(const V& v) {
    return g_capture ? v.data() == nullptr : v.data() != nullptr;
}

template <typename V>
static void expect(const V& v, U64 size, U64 refs) {
    CHECK(v.size() == size);
    CHECK(v.refs() == refs);
}

int main(int argc, char** argv) {
    g_capture = argc > 1 && std::strcmp(argv[1], "--capture") == 0;
    if (g_capture) Profiler::StartCapture();
    const U64 bytes = 24 * sizeof(U64);

    {   // empty
        ArrayTensor<Device::CUDA, U64> a;
        expect(a, 0, 0);
        CHECK(a.data() == nullptr);
        std::printf("empty creation\n");
    }
    {   // owning
        ArrayTensor<Device::CUDA, U64> a({1, 2, 3, 4});
        expect(a, 24, 1);
        CHECK(owns_memory(a));
        CHECK(a.size_bytes() == bytes);
        std::printf("filled creation\n");
    }
    {   // non-owning view of foreign memory: no count, never freed
        void* ptr = reinterpret_cast<void*>(0xdeadbeefc00feeull);
        ArrayTensor<Device::CUDA, U64> a(ptr, {1, 2, 3, 4});
        expect(a, 24, 0);
        CHECK(a.data() == ptr);
        ArrayTensor<Device::CUDA, U64> b(a);
        expect(b, 24, 0);
        CHECK(b.data() == ptr);
        std::printf("zombie creation\n");
    }
    {   // copy construction shares
        ArrayTensor<Device::CUDA, U64> a({1, 2, 3, 4});
        ArrayTensor<Device::CUDA, U64> c(a);
        expect(c, 24, 2);
        expect(a, 24, 2);
        CHECK(c.data() == a.data());
        CHECK(owns_memory(a));
        std::printf("copy constructor\n");
    }
    {   // move construction steals
        ArrayTensor<Device::CUDA, U64> a({1, 2, 3, 4});
        U64* const p = a.data();
        ArrayTensor<Device::CUDA, U64> c(std::move(a));
        expect(c, 24, 1);
        CHECK(c.data() == p);
        expect(a, 0, 0);
        CHECK(a.data() == nullptr);
        std::printf("move constructor\n");
    }
    {   // copy assignment shares
        ArrayTensor<Device::CUDA, U64> a({1, 2, 3, 4});
        ArrayTensor<Device::CUDA, U64> c;
        expect(c, 0, 0);
        c = a;
        expect(a, 24, 2);
        expect(c, 24, 2);
        CHECK(c.data() == a.data());
        c = c;   // self-assignment keeps the count
        expect(c, 24, 2);
        std::printf("copy assignment\n");
    }
    {   // move assignment steals
        ArrayTensor<Device::CUDA, U64> a({1, 2, 3, 4});
        ArrayTensor<Device::CUDA, U64> c;
        c = std::move(a);
        expect(a, 0, 0);
        CHECK(a.data() == nullptr);
        expect(c, 24, 1);
        CHECK(owns_memory(c));
        std::printf("move assignment\n");
    }
    {   // assignment over an owner releases the old memory
        ArrayTensor<Device::CUDA, U64> a({1, 2, 3, 4});
        ArrayTensor<Device::CUDA, U64> b({2, 2, 2, 2});
        ArrayTensor<Device::CUDA, U64> keep(b);
        expect(b, 16, 2);
        b = a;
        expect(a, 24, 2);
        expect(keep, 16, 1);
        std::printf("reassignment\n");
    }
    {   // aggregate member built in place
        struct Test {
            ArrayTensor<Device::CPU, U64> pickles;
        };
        Test test = Test{
            .pickles = ArrayTensor<Device::CPU, U64>({1, 2, 3, 4}),
        };
        expect(test.pickles, 24, 1);
        CHECK(owns_memory(test.pickles));
        std::printf("struct member\n");
    }
    {   // aggregate member copied from a named tensor
        struct Test {
            ArrayTensor<Device::CPU, U64> pickles;
        };
        auto pickles = ArrayTensor<Device::CPU, U64>({1, 2, 3, 4});
        Test test = Test{
            .pickles = pickles,
        };
        expect(test.pickles, 24, 2);
        CHECK(test.pickles.data() == pickles.data());
        expect(pickles, 24, 2);
        std::printf("initializer struct member\n");
    }
    if (g_capture) {
        // Every owner above has gone: the profiler saw each allocation freed.
        CHECK(Profiler::GetCudaCurrent() == 0);
        CHECK(Profiler::GetCpuCurrent() == 0);
        CHECK(Profiler::GetCudaAllocations() == 6 * bytes + 16 * sizeof(U64));
        CHECK(Profiler::GetCpuAllocations() == 2 * bytes);
        Profiler::StopCapture();
    }
    std::printf("vector semantics: %d failures\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
