// GUPPI RAW writer (reference include/blade/modules/guppi/writer.hh:14-99,
// src/modules/guppi/writer.cc:9-63): appends one block per process() call
// to <filepath>.0000.raw, with the header cards of the data shape (BLOCSIZE,
// OBSNCHAN, NANTS, NPOL, NBITS, DIRECTIO), NBEAM, DATATYPE = 'FLOAT' and,
// from the second block on, PKTIDX (time samples per block times the blocks
// written before the previous one, as the reference); headerPut adds or
// replaces cards.  Written without the guppirawc99 library (see reader.hh).
#ifndef BLADE_MODULES_GUPPI_WRITER_HH
#define BLADE_MODULES_GUPPI_WRITER_HH

#include <string>
#include <utility>
#include <vector>

#include "blade/module.hh"

namespace Blade::Modules::Guppi {

template <typename IT>
class BLADE_API Writer : public Module {
 public:
    struct Config {
        std::string filepath;
        bool directio = true;
        U64 inputFrequencyBatches;
        U64 blockSize = 512;
    };

    const Config& getConfig() const { return config; }

    struct Input {
        const ArrayTensor<Device::CPU, IT>& buf;
    };

    const ArrayTensor<Device::CPU, IT>& getInputBuffer() const { return input.buf; }

    struct Output {};

    Taint getTaint() const override { return Taint::CONSUMER; }
    std::string name() const override { return "Guppi Writer"; }

    explicit Writer(const Config& config, const Input& input, const Stream& stream = {});
    ~Writer() override;
    Result process(const U64& currentStepCount, const Stream& stream = {}) override;

    void headerPut(const std::string& key, const std::string& value);
    void headerPut(const std::string& key, const F64& value);
    void headerPut(const std::string& key, const I64& value);
    void headerPut(const std::string& key, const I32& value) { headerPut(key, static_cast<I64>(value)); }
    void headerPut(const std::string& key, const U64& value) { headerPut(key, static_cast<I64>(value)); }

 private:
    const Config config;
    const Input input;
    Output output;
    U64 writeCounter;
    int fileDescriptor;
    std::vector<std::pair<std::string, std::string>> cards;   // keyword, formatted value
};

}  // namespace Blade::Modules::Guppi

#endif
