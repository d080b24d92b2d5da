// Host-side code that runs on several threads or parses untrusted bytes,
// built without the HIP runtime being used, for ThreadSanitizer and
// AddressSanitizer + UBSan builds (tests/cpp/build_cpp.py builds it with g++;
// no device code is involved):
//
//   * the tensor reference count (include/blade/memory/vector.hh) under
//     copies and releases from many threads, with the allocation profiler
//     capturing (sizes recorded, nothing allocated);
//   * the GUPPI reader's thread pool (parallel_for) and its positioned reads
//     split over threads (read_full_parallel), checked byte for byte;
//   * the GUPPI header parser and file-sequence naming on well-formed and
//     malformed input.
//
// Usage: host_concurrency <scratch directory>.  Prints "host concurrency: N
// failures"; exit status 0 when every check holds.  A sanitizer report makes
// the sanitizer's own exit status non-zero.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "blade/memory/base.hh"
#include "guppi_io.hh"

namespace Blade {
// The profiler's singleton, which the library normally defines.
Profiler& Profiler::Get() {
    static Profiler p;
    return p;
}
}  // namespace Blade

using namespace Blade;
namespace gio = Blade::Modules::Guppi::io;

static int g_failures = 0;
#define CHECK(cond)                                                           \
    do {                                                                      \
        if (!(cond)) {                                                        \
            std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
            g_failures++;                                                     \
        }                                                                     \
    } while (0)

static void refcount_across_threads() {
    Profiler::StartCapture();
    {
        ArrayTensor<Device::CUDA, F32> a({2, 3, 4, 5});
        std::vector<std::thread> pool;
        std::atomic<int> go{0};
        for (int k = 0; k < 8; k++) {
            pool.emplace_back([&a, &go, k]() {
                while (go.load() == 0) {
                }
                ArrayTensor<Device::CUDA, F32> held;
                for (int i = 0; i < 20000; i++) {
                    ArrayTensor<Device::CUDA, F32> c(a);             // share
                    ArrayTensor<Device::CUDA, F32> m(std::move(c));   // steal
                    if ((i + k) % 3 == 0) held = m;                   // assign over
                    else if ((i + k) % 3 == 1) held = std::move(m);
                }
            });
        }
        go = 1;
        for (auto& t : pool) t.join();
        CHECK(a.refs() == 1);
        CHECK(Profiler::GetCudaCurrent() == a.size_bytes());
    }
    // the last owner released it exactly once
    CHECK(Profiler::GetCudaCurrent() == 0);
    CHECK(Profiler::GetCudaAllocations() == 2 * 3 * 4 * 5 * sizeof(F32));
    Profiler::StopCapture();
    std::printf("refcount across threads\n");
}

static void thread_pool() {
    for (unsigned threads : {1u, 2u, 7u, 16u}) {
        const size_t n = 1000;
        std::vector<std::atomic<int>> hits(n);
        for (auto& h : hits) h = 0;
        std::vector<size_t> out(n, 0);
        const bool ok = gio::parallel_for(n, threads, [&](size_t i) {
            hits[i]++;
            out[i] = i * i;
            return true;
        });
        CHECK(ok);
        bool once = true, right = true;
        for (size_t i = 0; i < n; i++) {
            once = once && hits[i].load() == 1;
            right = right && out[i] == i * i;
        }
        CHECK(once);
        CHECK(right);
        // a failing item stops the pool and is reported
        std::atomic<size_t> ran{0};
        const bool stopped = gio::parallel_for(n, threads, [&](size_t i) {
            ran++;
            return i != 500;
        });
        CHECK(!stopped);
        CHECK(ran.load() >= 501 && ran.load() <= n);
    }
    std::printf("thread pool\n");
}

static void parallel_reads(const std::string& dir) {
    const std::string path = dir + "/host_concurrency.bin";
    const size_t bytes = (37u << 20) + 123;
    std::vector<unsigned char> data(bytes);
    for (size_t i = 0; i < bytes; i++) data[i] = (unsigned char)((i * 2654435761u) >> 13);
    const int wfd = open(path.c_str(), O_CREAT | O_TRUNC | O_WRONLY, 0644);
    CHECK(wfd >= 0);
    CHECK(gio::write_full(wfd, data.data(), bytes));
    close(wfd);
    const int fd = open(path.c_str(), O_RDONLY);
    CHECK(fd >= 0);
    for (unsigned threads : {1u, 3u, 8u, 16u}) {
        std::vector<unsigned char> got(bytes, 0);
        CHECK(gio::read_full_parallel(fd, got.data(), bytes, 0, threads));
        CHECK(std::memcmp(got.data(), data.data(), bytes) == 0);
        // from an offset, to the end of the file
        const size_t off = 5000001, len = bytes - off;
        std::vector<unsigned char> tail(len, 0);
        CHECK(gio::read_full_parallel(fd, tail.data(), len, (off_t)off, threads));
        CHECK(std::memcmp(tail.data(), data.data() + off, len) == 0);
        // past the end fails instead of returning short data
        std::vector<unsigned char> over(1 << 20);
        CHECK(!gio::read_full_parallel(fd, over.data(), over.size(), (off_t)(bytes - 100), threads));
    }
    close(fd);
    unlink(path.c_str());
    std::printf("parallel reads\n");
}

static std::string card(const std::string& key, const std::string& value) {
    std::string c = key;
    c.resize(8, ' ');
    c += "= " + value;
    c.resize(80, ' ');
    return c;
}

static void header_parser(const std::string& dir) {
    const std::string path = dir + "/host_concurrency.raw";
    auto write = [&](const std::string& text) {
        const int fd = open(path.c_str(), O_CREAT | O_TRUNC | O_WRONLY, 0644);
        CHECK(fd >= 0 && gio::write_full(fd, text.data(), text.size()));
        close(fd);
    };
    auto parse = [&](gio::Block& blk, off_t& data_offset, size_t& blocsize) -> int {
        const int fd = open(path.c_str(), O_RDONLY);
        struct stat st;
        fstat(fd, &st);
        int r;
        try {
            r = gio::parse_header(fd, 0, st.st_size, blk, data_offset, blocsize) ? 1 : 0;
        } catch (const std::runtime_error&) {
            r = -1;
        }
        close(fd);
        return r;
    };
    gio::Block blk;
    off_t off = 0;
    size_t bs = 0;
    // well-formed, no DIRECTIO
    std::string h = card("BLOCSIZE", "1000") + card("SRC_NAME", "'  3C286 '") + card("NBITS", "8");
    h += std::string("END").append(77, ' ');
    write(h + std::string(1000, 'x'));
    CHECK(parse(blk, off, bs) == 1);
    CHECK(off == 4 * 80 && bs == 1000);
    CHECK(blk.cards["SRC_NAME"] == "3C286");
    CHECK(blk.cards["NBITS"] == "8");
    // DIRECTIO pads header and data to 512 bytes
    h = card("BLOCSIZE", "1000") + card("DIRECTIO", "1") + std::string("END").append(77, ' ');
    write(h + std::string(2000, 'x'));
    CHECK(parse(blk, off, bs) == 1);
    CHECK(off == 512 && bs == 1024);
    // no END card before the end of the file
    write(card("BLOCSIZE", "1000") + card("NBITS", "8"));
    CHECK(parse(blk, off, bs) == -1);
    // no BLOCSIZE
    write(card("NBITS", "8") + std::string("END").append(77, ' '));
    CHECK(parse(blk, off, bs) == -1);
    // a partial card at the end, and an empty file: clean end of file
    write(std::string(40, ' '));
    CHECK(parse(blk, off, bs) == 0);
    write("");
    CHECK(parse(blk, off, bs) == 0);
    // cards without '=' in column 9 are skipped; END with trailing text is not END
    h = card("BLOCSIZE", "64") + std::string("COMMENT no equals sign").append(58, ' ') +
        std::string("ENDX").append(76, ' ') + std::string("END").append(77, ' ');
    write(h + std::string(64, 'x'));
    CHECK(parse(blk, off, bs) == 1);
    CHECK(off == 4 * 80 && bs == 64 && blk.cards.size() == 1);
    unlink(path.c_str());

    std::string stem;
    int index = -1;
    CHECK(gio::split_stem("obs/guppi_1.0007.raw", stem, index) && stem == "obs/guppi_1" && index == 7);
    CHECK(gio::split_stem(".0000.raw", stem, index) && stem.empty() && index == 0);
    CHECK(!gio::split_stem("0000.raw", stem, index));
    CHECK(!gio::split_stem("guppi.00a1.raw", stem, index));
    CHECK(!gio::split_stem("guppi.0001.fil", stem, index));
    CHECK(!gio::split_stem("", stem, index));
    std::printf("header parser\n");
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    refcount_across_threads();
    thread_pool();
    parallel_reads(dir);
    header_parser(dir);
    std::printf("host concurrency: %d failures\n", g_failures);
    return g_failures == 0 ? 0 : 1;
}
