// File-level helpers of the GUPPI RAW reader and writer: positioned reads
// split over a pool of host threads, header card parsing and the
// "<stem>.NNNN.raw" sequence naming.  Header-only so that the host
// concurrency and parsing code can also be built on its own (the sanitizer
// tests under tests/cpp/sanitize) without the HIP runtime.
#ifndef BLADE_AMD_HOST_GUPPI_IO_HH
#define BLADE_AMD_HOST_GUPPI_IO_HH

#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace Blade::Modules::Guppi::io {

constexpr size_t kCard = 80;
constexpr size_t kDirectIoAlign = 512;

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

inline bool read_full(int fd, void* dst, size_t bytes, off_t offset) {
    unsigned char* p = static_cast<unsigned char*>(dst);
    while (bytes > 0) {
        const ssize_t r = pread(fd, p, bytes, offset);
        if (r <= 0) return false;
        p += r;
        offset += r;
        bytes -= (size_t)r;
    }
    return true;
}

// Host threads for the reads of one step: a single thread copies from the
// page cache at ~20 GB/s, several together at several times that.
// BLADE_GUPPI_READ_THREADS overrides the default of 8 (1 = serial).
inline unsigned read_threads() {
    const char* v = getenv("BLADE_GUPPI_READ_THREADS");
    const int n = v ? atoi(v) : 8;
    return (unsigned)std::max(1, std::min(n, 64));
}

// Runs fn(i) for i in [0, n) on up to `threads` threads; true if every call
// returned true.
template <typename Fn>
bool parallel_for(size_t n, unsigned threads, Fn fn) {
    threads = (unsigned)std::min<size_t>(threads, n);
    if (threads <= 1) {
        for (size_t i = 0; i < n; i++)
            if (!fn(i)) return false;
        return true;
    }
    std::atomic<size_t> next{0};
    std::atomic<bool> ok{true};
    auto worker = [&]() {
        for (size_t i = next++; i < n && ok.load(std::memory_order_relaxed); i = next++)
            if (!fn(i)) ok = false;
    };
    std::vector<std::thread> pool;
    pool.reserve(threads - 1);
    for (unsigned k = 1; k < threads; k++) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
    return ok.load();
}

// read_full over `threads` threads in pieces of at least 4 MiB.
inline bool read_full_parallel(int fd, void* dst, size_t bytes, off_t offset, unsigned threads) {
    constexpr size_t kPiece = 4u << 20;
    const size_t n = std::max<size_t>(1, std::min<size_t>(threads * 4, bytes / kPiece));
    const size_t step = (bytes + n - 1) / n;
    unsigned char* p = static_cast<unsigned char*>(dst);
    return parallel_for(n, threads, [&](size_t i) {
        const size_t b = i * step, e = std::min(bytes, b + step);
        return b >= e || read_full(fd, p + b, e - b, offset + (off_t)b);
    });
}

inline bool write_full(int fd, const void* src, size_t bytes) {
    const unsigned char* p = static_cast<const unsigned char*>(src);
    while (bytes > 0) {
        const ssize_t r = write(fd, p, bytes);
        if (r <= 0) return false;
        p += r;
        bytes -= (size_t)r;
    }
    return true;
}

inline std::string trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && (s[b] == ' ' || s[b] == '\'')) b++;
    while (e > b && (s[e - 1] == ' ' || s[e - 1] == '\'')) e--;
    return s.substr(b, e - b);
}

// One block of one file: its header cards and where its data starts.
struct Block {
    int file;
    off_t data_offset;
    std::map<std::string, std::string> cards;
};

// Parses the header at `offset`; returns false at a clean end of file.
inline bool parse_header(int fd, off_t offset, off_t file_size, Block& blk, off_t& data_offset, size_t& blocsize) {
    if (offset + (off_t)kCard > file_size) return false;
    blk.cards.clear();
    char card[kCard + 1];
    card[kCard] = 0;
    off_t pos = offset;
    for (;;) {
        if (pos + (off_t)kCard > file_size || !read_full(fd, card, kCard, pos))
            throw std::runtime_error("GUPPI header without END card");
        pos += kCard;
        if (std::strncmp(card, "END", 3) == 0 && (card[3] == ' ' || card[3] == 0)) break;
        const std::string line(card, kCard);
        if (line.size() < 10 || line[8] != '=') continue;
        blk.cards[trim(line.substr(0, 8))] = trim(line.substr(9));
    }
    const auto dio = blk.cards.find("DIRECTIO");
    const bool directio = dio != blk.cards.end() && std::atoi(dio->second.c_str()) != 0;
    data_offset = directio ? (off_t)align_up((size_t)(pos - offset), kDirectIoAlign) + offset : pos;
    const auto bs = blk.cards.find("BLOCSIZE");
    if (bs == blk.cards.end()) throw std::runtime_error("GUPPI header without BLOCSIZE");
    blocsize = std::strtoull(bs->second.c_str(), nullptr, 10);
    if (directio) blocsize = align_up(blocsize, kDirectIoAlign);
    return true;
}

// "<stem>.NNNN.raw" -> ("<stem>", NNNN); anything else is a single file.
inline bool split_stem(const std::string& path, std::string& stem, int& index) {
    const size_t n = path.size();
    if (n < 9 || path.compare(n - 4, 4, ".raw") != 0 || path[n - 9] != '.') return false;
    for (size_t i = n - 8; i < n - 4; i++)
        if (path[i] < '0' || path[i] > '9') return false;
    stem = path.substr(0, n - 9);
    index = std::atoi(path.substr(n - 8, 4).c_str());
    return true;
}

}  // namespace Blade::Modules::Guppi::io

#endif
