// Phasor module interface (reference include/blade/modules/phasor/generic.hh:14-58).
#ifndef BLADE_MODULES_PHASOR_GENERIC_HH
#define BLADE_MODULES_PHASOR_GENERIC_HH

#include <string>
#include <vector>

#include "blade/module.hh"
#include "blade/types.hh"

namespace Blade::Modules::Phasor {

template <typename OT>
class BLADE_API Generic : public Module {
 public:
    struct Config {
        U64 numberOfAntennas;
        U64 numberOfFrequencyChannels;
        U64 numberOfPolarizations;

        F64 observationFrequencyHz;
        F64 channelBandwidthHz;
        F64 totalBandwidthHz;
        U64 frequencyStartIndex;

        U64 referenceAntennaIndex;
        LLA arrayReferencePosition;
        RA_DEC boresightCoordinate;
        std::vector<XYZ> antennaPositions;
        ArrayTensor<Device::CPU, CF64> antennaCalibrations;
        std::vector<RA_DEC> beamCoordinates;

        U64 blockSize = 512;
    };

    const Config& getConfig() const { return config; }

    struct Input {
        const Tensor<Device::CPU, F64>& blockJulianDate;
        const Tensor<Device::CPU, F64>& blockDut1;
    };

    const Tensor<Device::CPU, F64>& getInputJulianDate() const { return input.blockJulianDate; }
    const Tensor<Device::CPU, F64>& getInputDut1() const { return input.blockDut1; }

    struct Output {
        DelayTensor<Device::CPU, F64> delays;
        PhasorTensor<Device::CUDA, OT> phasors;
    };

    const DelayTensor<Device::CPU, F64>& getOutputDelays() const { return output.delays; }
    const PhasorTensor<Device::CUDA, OT>& getOutputPhasors() const { return output.phasors; }

    Taint getTaint() const override { return Taint::CONSUMER | Taint::PRODUCER; }
    std::string name() const override { return "Phasor"; }
    // The delays are computed on the host for each block's Julian date.
    bool graphCapturable() const override { return false; }

    explicit Generic(const Config& config, const Input& input, const Stream& stream = {});
    virtual ~Generic() = default;
    Result process(const U64& currentStepCount, const Stream& stream = {}) override = 0;

 protected:
    const Config config;
    const Input input;
    Output output;

    U64 numberOfFrequencySteps;
    U64 currentFrequencyStep;
};

}  // namespace Blade::Modules::Phasor

#endif
