// GUPPI RAW reader (reference include/blade/modules/guppi/reader.hh:16-129,
// src/modules/guppi/reader.cc:50-168).  The reference reads through the
// guppirawc99 library (an un-vendored submodule); this module parses the
// format itself: a file is a sequence of blocks, each an ASCII header of
// 80-byte "KEYWORD = value" cards closed by an END card (padded to a
// multiple of 512 bytes when DIRECTIO = 1) followed by BLOCSIZE bytes of
// complex 8-bit samples laid out [aspect][channel][time][pol][re, im].
// Files continue as <stem>.0001.raw, <stem>.0002.raw, ...
#ifndef BLADE_MODULES_GUPPI_READER_HH
#define BLADE_MODULES_GUPPI_READER_HH

#include <memory>
#include <string>

#include "blade/module.hh"

namespace Blade::Modules::Guppi {

template <typename OT>
class BLADE_API Reader : public Module {
 public:
    struct Config {
        std::string filepath;
        U64 stepNumberOfTimeSamples;
        U64 stepNumberOfFrequencyChannels;
        U64 stepNumberOfAspects;
        U64 blockSize = 512;
    };

    const Config& getConfig() const { return config; }

    struct Input {};

    struct Output {
        ArrayTensor<Device::CPU, OT> stepBuffer;
        Tensor<Device::CPU, F64> stepJulianDate;
        Tensor<Device::CPU, F64> stepDut1;
    };

    const ArrayTensor<Device::CPU, OT>& getStepOutputBuffer() const { return output.stepBuffer; }
    const Tensor<Device::CPU, F64>& getStepOutputJulianDate() const { return output.stepJulianDate; }
    const Tensor<Device::CPU, F64>& getStepOutputDut1() const { return output.stepDut1; }

    ArrayShape getTotalOutputBufferShape() const;
    ArrayShape getStepOutputBufferShape() const;
    U64 getNumberOfSteps() const;

    Taint getTaint() const override { return Taint::PRODUCER; }
    std::string name() const override { return "Guppi Reader"; }

    explicit Reader(const Config& config, const Input& input, const Stream& stream = {});
    ~Reader() override;
    Result process(const U64& currentStepCount, const Stream& stream = {}) override;

    F64 getTotalBandwidth() const;
    F64 getChannelBandwidth() const;
    U64 getChannelStartIndex() const;
    F64 getObservationFrequency() const;

    struct State;

 private:
    Config config;
    const Input input;
    Output output;
    std::unique_ptr<State> state;
};

}  // namespace Blade::Modules::Guppi

#endif
