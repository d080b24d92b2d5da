// GUPPI RAW reader and writer (reference src/modules/guppi/reader.cc:50-168,
// src/modules/guppi/writer.cc:9-63), with the format parsed and written here
// rather than through the guppirawc99 library the reference links.
//
// Block = header + data.  Header: 80-byte ASCII cards "KEYWORD = value"
// (keyword left-justified in 8 columns, '=' in column 9, strings quoted with
// '), closed by a card that starts with "END"; with DIRECTIO = 1 the header
// and the data are each padded to a multiple of 512 bytes.  Data: BLOCSIZE
// bytes, [NANTS][OBSNCHAN/NANTS][time][pol][re, im] at NBITS per component.
// NPOL = 4 in a voltage header means two complex polarizations.
#define BL_LOG_DOMAIN "M::GUPPI"

#include "blade/modules/guppi/base.hh"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <map>
#include <stdexcept>
#include <thread>
#include <vector>

#include "blade/logger.hh"
#include "guppi_io.hh"

namespace Blade::Modules::Guppi {

using namespace io;

template <typename OT>
struct Reader<OT>::State {
    std::vector<int> fds;
    std::vector<Block> blocks;
    U64 n_aspect = 1, n_aspectchan = 0, n_time = 0, n_pol = 0, n_bit = 0;
    U64 t0 = 0, f0 = 0, a0 = 0;         // next step's first sample, channel, aspect
    bool exhausted = false;

    F64 num(const Block& b, const char* key, F64 dflt) const {
        const auto it = b.cards.find(key);
        return it == b.cards.end() ? dflt : std::strtod(it->second.c_str(), nullptr);
    }

    ~State() {
        for (int fd : fds) close(fd);
    }
};

template <typename OT>
Reader<OT>::Reader(const Config& config, const Input& input, const Stream&)
    : config(config), input(input), state(std::make_unique<State>()) {
    if (config.filepath.empty()) {
        BL_FATAL("Input file ({}) is invalid.", config.filepath);
        throw Result::ASSERTION_ERROR;
    }
    std::vector<std::string> paths;
    std::string stem;
    int index = 0;
    if (split_stem(config.filepath, stem, index)) {
        for (int i = index; i < 10000; i++) {
            char name[16];
            std::snprintf(name, sizeof(name), ".%04d.raw", i);
            const std::string p = stem + name;
            struct stat st;
            if (stat(p.c_str(), &st) != 0) break;
            paths.push_back(p);
        }
    } else {
        paths.push_back(config.filepath);
    }
    try {
        for (const auto& p : paths) {
            const int fd = open(p.c_str(), O_RDONLY);
            if (fd < 0) {
                BL_FATAL("Errored opening {}: {}.", p, std::strerror(errno));
                throw Result::ASSERTION_ERROR;
            }
            state->fds.push_back(fd);
            struct stat st;
            fstat(fd, &st);
            off_t pos = 0;
            for (;;) {
                Block b;
                b.file = (int)state->fds.size() - 1;
                off_t data = 0;
                size_t bytes = 0;
                if (!parse_header(fd, pos, st.st_size, b, data, bytes)) break;
                if (data + (off_t)bytes > st.st_size) break;     // truncated last block
                b.data_offset = data;
                state->blocks.push_back(std::move(b));
                pos = data + (off_t)bytes;
            }
        }
    } catch (const std::runtime_error& e) {
        BL_FATAL("GUPPI parse error in {}: {}.", config.filepath, e.what());
        throw Result::ASSERTION_ERROR;
    }
    if (state->blocks.empty()) {
        BL_FATAL("No GUPPI RAW block found in {}.", config.filepath);
        throw Result::ASSERTION_ERROR;
    }
    const Block& b0 = state->blocks.front();
    const U64 obsnchan = (U64)state->num(b0, "OBSNCHAN", 0);
    const U64 npol = (U64)state->num(b0, "NPOL", 2);
    state->n_bit = (U64)state->num(b0, "NBITS", 8);
    state->n_aspect = (U64)state->num(b0, "NANTS", 1);
    state->n_pol = npol == 4 ? 2 : npol;
    if (state->n_bit != 8 || sizeof(OT) != 2) {
        BL_FATAL("Only 8-bit complex samples are supported (NBITS = {}).", state->n_bit);
        throw Result::ASSERTION_ERROR;
    }
    if (obsnchan == 0 || state->n_aspect == 0 || obsnchan % state->n_aspect) {
        BL_FATAL("OBSNCHAN ({}) must be a positive multiple of NANTS ({}).", obsnchan, state->n_aspect);
        throw Result::ASSERTION_ERROR;
    }
    state->n_aspectchan = obsnchan / state->n_aspect;
    const U64 blocsize = (U64)state->num(b0, "BLOCSIZE", 0);
    const U64 per_time = obsnchan * state->n_pol * 2 * state->n_bit / 8;
    if (blocsize == 0 || blocsize % per_time) {
        BL_FATAL("BLOCSIZE ({}) is not a whole number of time samples.", blocsize);
        throw Result::ASSERTION_ERROR;
    }
    state->n_time = blocsize / per_time;

    if (this->config.stepNumberOfAspects == 0) this->config.stepNumberOfAspects = state->n_aspect;
    if (this->config.stepNumberOfFrequencyChannels == 0)
        this->config.stepNumberOfFrequencyChannels = state->n_aspectchan;
    if (this->config.stepNumberOfTimeSamples == 0) this->config.stepNumberOfTimeSamples = state->n_time;
    const ArrayShape total = getTotalOutputBufferShape();
    if (state->n_aspect % this->config.stepNumberOfAspects ||
        state->n_aspectchan % this->config.stepNumberOfFrequencyChannels ||
        this->config.stepNumberOfTimeSamples > total.numberOfTimeSamples()) {
        BL_FATAL("Step shape {} does not tile the file's shape {}.", getStepOutputBufferShape(), total);
        throw Result::ASSERTION_ERROR;
    }

    output.stepDut1 = Tensor<Device::CPU, F64>({1});
    output.stepJulianDate = Tensor<Device::CPU, F64>({1});
    output.stepBuffer = ArrayTensor<Device::CPU, OT>(getStepOutputBufferShape());

    BL_INFO("Type: {} -> {}", "N/A", TypeInfo<OT>::name);
    BL_INFO("Step Shape: {} -> {}", "N/A", getStepOutputBuffer().shape());
    BL_INFO("Total Shape: {} -> {}", "N/A", total);
    BL_INFO("Input File Path: {}", config.filepath);
}

template <typename OT>
Reader<OT>::~Reader() = default;

template <typename OT>
ArrayShape Reader<OT>::getTotalOutputBufferShape() const {
    return ArrayShape({state->n_aspect, state->n_aspectchan, state->n_time * state->blocks.size(), state->n_pol});
}

template <typename OT>
ArrayShape Reader<OT>::getStepOutputBufferShape() const {
    return ArrayShape({config.stepNumberOfAspects, config.stepNumberOfFrequencyChannels,
                       config.stepNumberOfTimeSamples, state->n_pol});
}

template <typename OT>
U64 Reader<OT>::getNumberOfSteps() const {
    return getTotalOutputBufferShape().size() / getStepOutputBufferShape().size();
}

template <typename OT>
F64 Reader<OT>::getChannelBandwidth() const {
    return state->num(state->blocks.front(), "CHAN_BW", 0.0) * 1e6;
}

template <typename OT>
F64 Reader<OT>::getTotalBandwidth() const {
    return getChannelBandwidth() * getStepOutputBufferShape().numberOfFrequencyChannels();
}

template <typename OT>
U64 Reader<OT>::getChannelStartIndex() const {
    return (U64)state->num(state->blocks.front(), "SCHAN", 0.0);
}

template <typename OT>
F64 Reader<OT>::getObservationFrequency() const {
    return state->num(state->blocks.front(), "OBSFREQ", 0.0) * 1e6;
}

// Steps walk time first, then channels, then aspects; a step needs the
// step's time samples left on the time axis (reference keepRunning()).
template <typename OT>
Result Reader<OT>::process(const U64&, const Stream&) {
    State& st = *state;
    const U64 nA = config.stepNumberOfAspects, nF = config.stepNumberOfFrequencyChannels;
    const U64 nT = config.stepNumberOfTimeSamples;
    const U64 total_t = st.n_time * st.blocks.size();
    if (st.exhausted || st.t0 + nT > total_t) return Result::PIPELINE_EXHAUSTED;

    const U64 block_index = st.t0 / st.n_time;
    const Block& bm = st.blocks[block_index];
    // Time of the step (reference reader.cc:128-139): the block's PKTIDX plus
    // (block index + PIPERBLK / 2) * samples per block, as an integer packet
    // index, at TBIN * samples per block / PIPERBLK seconds per packet after
    // SYNCTIME (the guppirawc99 date routine, restated); Julian date =
    // unix / 86400 + 2440587.5.
    const F64 tbin = 1.0 / getChannelBandwidth();
    U64 piperblk = (U64)st.num(bm, "PIPERBLK", 0.0);
    if (piperblk == 0) piperblk = st.n_time;
    const U64 synctime = (U64)st.num(bm, "SYNCTIME", 0.0);
    const U64 pktidx = (U64)st.num(bm, "PKTIDX", 0.0) +
                       (U64)(((F64)block_index + 0.5 * (F64)piperblk) * (F64)st.n_time);
    const F64 unix = (F64)synctime + (F64)pktidx * ((tbin * (F64)st.n_time) / (F64)piperblk);
    output.stepJulianDate.data()[0] = unix / 86400.0 + 2440587.5;
    output.stepDut1.data()[0] = st.num(bm, "DUT1", 0.0);

    // Copy [a0, a0+nA) x [f0, f0+nF) x [t0, t0+nT) row by row; a step that is
    // one whole block at full width is one contiguous read.
    const size_t sample = st.n_pol * 2;                      // bytes per time sample
    unsigned char* dst = reinterpret_cast<unsigned char*>(output.stepBuffer.data());
    const bool whole = nA == st.n_aspect && nF == st.n_aspectchan && nT == st.n_time && st.t0 % st.n_time == 0;
    // Rows (and pieces of a whole block) are read by several host threads.
    const unsigned threads = read_threads();
    if (whole) {
        if (!read_full_parallel(st.fds[bm.file], dst, nA * nF * nT * sample, bm.data_offset, threads))
            return Result::ERROR;
    } else {
        const bool ok = parallel_for(nA * nF, threads, [&](size_t row) {
            const U64 a = row / nF, f = row % nF;
            U64 t = st.t0, done = 0;
            while (done < nT) {
                const U64 b = t / st.n_time, tb = t % st.n_time;
                const U64 run = std::min(nT - done, st.n_time - tb);
                const Block& blk = st.blocks[b];
                const off_t off = blk.data_offset +
                    (off_t)((((st.a0 + a) * st.n_aspectchan + (st.f0 + f)) * st.n_time + tb) * sample);
                if (!read_full(st.fds[blk.file], dst + ((a * nF + f) * nT + done) * sample, run * sample, off))
                    return false;
                done += run;
                t += run;
            }
            return true;
        });
        if (!ok) return Result::ERROR;
    }

    st.t0 += nT;
    if (st.t0 + nT > total_t) {
        st.t0 = 0;
        st.f0 += nF;
        if (st.f0 >= st.n_aspectchan) {
            st.f0 = 0;
            st.a0 += nA;
            if (st.a0 >= st.n_aspect) st.exhausted = true;
        }
    }
    return Result::SUCCESS;
}

template class BLADE_API Reader<CI8>;

// ---------------------------------------------------------------- writer

namespace {

std::string format_card(const std::string& key, const std::string& value) {
    char card[kCard + 1];
    std::snprintf(card, sizeof(card), "%-8.8s= %-70.70s", key.c_str(), value.c_str());
    return std::string(card, kCard);
}

}  // namespace

template <typename IT>
Writer<IT>::Writer(const Config& config, const Input& input, const Stream&)
    : config(config), input(input), writeCounter(0), fileDescriptor(-1) {
    if (config.inputFrequencyBatches != 1) {
        BL_FATAL("Frequency-batched input ({} batches) is not supported; use 1.", config.inputFrequencyBatches);
        throw Result::ERROR;
    }
    char name[16];
    std::snprintf(name, sizeof(name), ".%04d.raw", 0);
    const std::string path = config.filepath + name;
    fileDescriptor = open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fileDescriptor < 0) {
        BL_FATAL("Could not open '{}': {}.", path, std::strerror(errno));
        throw Result::ERROR;
    }
    const ArrayShape& s = getInputBuffer().shape();
    const U64 nbit = sizeof(IT) * 8 / 2;
    headerPut("BLOCSIZE", (I64)(s.size() * sizeof(IT)));
    headerPut("OBSNCHAN", (I64)(s.numberOfAspects() * s.numberOfFrequencyChannels()));
    headerPut("NANTS", (I64)s.numberOfAspects());
    headerPut("NPOL", (I64)s.numberOfPolarizations());
    headerPut("NBITS", (I64)nbit);
    headerPut("DIRECTIO", (I64)(config.directio ? 1 : 0));
    headerPut("NBEAM", (I64)s.numberOfAspects());
    headerPut("DATATYPE", std::string("FLOAT"));
    BL_INFO("Type: {} -> {}", TypeInfo<IT>::name, "N/A");
    BL_INFO("Shape: {} -> {}", getInputBuffer().shape(), "N/A");
    BL_INFO("Output File Path: {}", path);
    BL_INFO("Direct I/O: {}", config.directio ? "YES" : "NO");
}

template <typename IT>
Writer<IT>::~Writer() {
    if (fileDescriptor >= 0) close(fileDescriptor);
}

template <typename IT>
void Writer<IT>::headerPut(const std::string& key, const std::string& value) {
    const std::string v = "'" + value + "'";
    for (auto& c : cards)
        if (c.first == key) {
            c.second = v;
            return;
        }
    cards.emplace_back(key, v);
}

template <typename IT>
void Writer<IT>::headerPut(const std::string& key, const F64& value) {
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%.17g", value);
    for (auto& c : cards)
        if (c.first == key) {
            c.second = buf;
            return;
        }
    cards.emplace_back(key, buf);
}

template <typename IT>
void Writer<IT>::headerPut(const std::string& key, const I64& value) {
    const std::string v = std::to_string(value);
    for (auto& c : cards)
        if (c.first == key) {
            c.second = v;
            return;
        }
    cards.emplace_back(key, v);
}

template <typename IT>
Result Writer<IT>::process(const U64&, const Stream&) {
    std::string header;
    for (const auto& c : cards) header += format_card(c.first, c.second);
    char end[kCard + 1];
    std::snprintf(end, sizeof(end), "%-80s", "END");
    header += std::string(end, kCard);
    const size_t data_bytes = getInputBuffer().size() * sizeof(IT);
    if (config.directio) header.resize(align_up(header.size(), kDirectIoAlign), ' ');
    if (!write_full(fileDescriptor, header.data(), header.size())) return Result::ERROR;
    if (!write_full(fileDescriptor, getInputBuffer().data(), data_bytes)) return Result::ERROR;
    if (config.directio && data_bytes % kDirectIoAlign) {
        const std::vector<char> pad(align_up(data_bytes, kDirectIoAlign) - data_bytes, 0);
        if (!write_full(fileDescriptor, pad.data(), pad.size())) return Result::ERROR;
    }
    // As the reference (writer.cc:59): the card written with block k + 1 is
    // PKTIDX = (k - 1) * time samples, so the first two blocks carry none and 0.
    headerPut("PKTIDX", (I64)(getInputBuffer().shape().numberOfTimeSamples() * writeCounter++));
    return Result::SUCCESS;
}

template class BLADE_API Writer<CF16>;
template class BLADE_API Writer<CF32>;

}  // namespace Blade::Modules::Guppi
