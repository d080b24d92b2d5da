// ATA Mode-B bundle (reference include/blade/bundles/ata/mode_b.hh:17-225):
// Cast -> Channelizer -> Polarizer -> Phasor -> Beamformer -> [Detector -> Cast].
// With CI8 input, rate 1 and no circular conversion the first three stages
// are identities on the data, so the beamformer reads the CI8 voltages
// directly (the widening happens in its loads) and no CF32 copy of the input
// is ever written.
#ifndef BLADE_BUNDLES_ATA_MODE_B_HH
#define BLADE_BUNDLES_ATA_MODE_B_HH

#include <memory>
#include <type_traits>
#include <vector>

#include "blade/bundle.hh"
#include "blade/modules/base.hh"

namespace Blade::Bundles::ATA {

template <typename IT, typename OT>
class BLADE_API ModeB : public Bundle {
 public:
    struct Config {
        ArrayShape inputShape;
        ArrayShape outputShape;

        U64 preBeamformerChannelizerRate;
        BOOL preBeamformerPolarizerConvertToCircular = false;

        F64 phasorObservationFrequencyHz;
        F64 phasorChannelBandwidthHz;
        F64 phasorTotalBandwidthHz;
        U64 phasorFrequencyStartIndex;
        U64 phasorReferenceAntennaIndex;
        LLA phasorArrayReferencePosition;
        RA_DEC phasorBoresightCoordinate;
        std::vector<XYZ> phasorAntennaPositions;
        ArrayTensor<Device::CPU, CF64> phasorAntennaCalibrations;
        std::vector<RA_DEC> phasorBeamCoordinates;

        BOOL beamformerIncoherentBeam = false;

        BOOL detectorEnable = false;
        U64 detectorIntegrationSize;
        U64 detectorNumberOfOutputPolarizations;

        U64 castBlockSize = 512;
        U64 channelizerBlockSize = 512;
        U64 phasorBlockSize = 512;
        U64 beamformerBlockSize = 512;
        U64 polarizerBlockSize = 512;
        U64 detectorBlockSize = 512;

        BOOL beamformerSplitFp16 = false;
    };

    const Config& getConfig() const { return config; }

    struct Input {
        const Tensor<Device::CPU, F64>& dut;
        const Tensor<Device::CPU, F64>& julianDate;
        const ArrayTensor<Device::CUDA, IT>& buffer;
    };

    const ArrayTensor<Device::CUDA, IT>& getInputBuffer() const { return input.buffer; }
    const Tensor<Device::CPU, F64>& getInputDut() const { return input.dut; }
    const Tensor<Device::CPU, F64>& getInputJulianDate() const { return input.julianDate; }

    const ArrayTensor<Device::CUDA, OT>& getOutputBuffer() {
        if (config.detectorEnable) {
            if constexpr (!std::is_same<OT, F32>::value) return outputCast->getOutputBuffer();
            else return detector->getOutputBuffer();
        } else {
            if constexpr (!std::is_same<OT, CF32>::value) return complexOutputCast->getOutputBuffer();
            else return beamformerOutput();
        }
    }

    explicit ModeB(const Config& config, const Input& input, const Stream& stream)
        : Bundle(stream), config(config), input(input) {
        const bool direct = std::is_same<IT, CI8>::value && config.preBeamformerChannelizerRate == 1 &&
                            !config.preBeamformerPolarizerConvertToCircular;
        const ArrayShape* bfShape = nullptr;
        if (!direct) {
            this->connect(inputCast, {.blockSize = config.castBlockSize}, {.buf = input.buffer});
            this->connect(channelizer, {.rate = config.preBeamformerChannelizerRate,
                                        .blockSize = config.channelizerBlockSize},
                          {.buf = inputCast->getOutputBuffer()});
            this->connect(polarizer, {.inputPolarization = POL::XY,
                                      .outputPolarization = config.preBeamformerPolarizerConvertToCircular
                                                                ? POL::LR : POL::XY,
                                      .blockSize = config.polarizerBlockSize},
                          {.buf = channelizer->getOutputBuffer()});
            bfShape = &polarizer->getOutputBuffer().shape();
        } else {
            bfShape = &input.buffer.shape();
        }
        this->connect(phasor, {
            .numberOfAntennas = bfShape->numberOfAspects(),
            .numberOfFrequencyChannels = bfShape->numberOfFrequencyChannels(),
            .numberOfPolarizations = bfShape->numberOfPolarizations(),
            .observationFrequencyHz = config.phasorObservationFrequencyHz,
            .channelBandwidthHz = config.phasorChannelBandwidthHz,
            .totalBandwidthHz = config.phasorTotalBandwidthHz,
            .frequencyStartIndex = config.phasorFrequencyStartIndex,
            .referenceAntennaIndex = config.phasorReferenceAntennaIndex,
            .arrayReferencePosition = config.phasorArrayReferencePosition,
            .boresightCoordinate = config.phasorBoresightCoordinate,
            .antennaPositions = config.phasorAntennaPositions,
            .antennaCalibrations = config.phasorAntennaCalibrations,
            .beamCoordinates = config.phasorBeamCoordinates,
            .blockSize = config.phasorBlockSize,
        }, {.blockJulianDate = input.julianDate, .blockDut1 = input.dut});
        const typename Modules::Beamformer::Generic<CF32, CF32>::Config bfc = {
            .enableIncoherentBeam = config.beamformerIncoherentBeam,
            .enableIncoherentBeamSqrt = config.detectorEnable,
            .blockSize = config.beamformerBlockSize,
            .enableSplitFp16 = config.beamformerSplitFp16 && direct,
        };
        if (direct) {
            if constexpr (std::is_same<IT, CI8>::value) {
                this->connect(beamformerCi8, {bfc.enableIncoherentBeam, bfc.enableIncoherentBeamSqrt,
                                              bfc.blockSize, bfc.enableSplitFp16},
                              {.buf = input.buffer, .phasors = phasor->getOutputPhasors()});
            }
        } else {
            this->connect(beamformer, bfc, {.buf = polarizer->getOutputBuffer(),
                                            .phasors = phasor->getOutputPhasors()});
        }
        if (config.detectorEnable) {
            this->connect(detector, {.integrationSize = config.detectorIntegrationSize,
                                     .numberOfOutputPolarizations = config.detectorNumberOfOutputPolarizations,
                                     .blockSize = config.detectorBlockSize},
                          {.buf = beamformerOutput()});
            if constexpr (!std::is_same<OT, F32>::value) {
                this->connect(outputCast, {.blockSize = config.castBlockSize},
                              {.buf = detector->getOutputBuffer()});
            }
        } else {
            if constexpr (!std::is_same<OT, CF32>::value) {
                this->connect(complexOutputCast, {.blockSize = config.castBlockSize},
                              {.buf = beamformerOutput()});
            }
        }
        if (getOutputBuffer().shape() != config.outputShape) {
            BL_FATAL("Expected output buffer size ({}) mismatch with actual size ({}).",
                     config.outputShape, getOutputBuffer().shape());
            throw Result::ERROR;
        }
    }

 private:
    const Config config;
    Input input;

    std::shared_ptr<Modules::Cast<IT, CF32>> inputCast;
    std::shared_ptr<Modules::Channelizer<CF32, CF32>> channelizer;
    std::shared_ptr<Modules::Phasor::ATA<CF32>> phasor;
    std::shared_ptr<Modules::Beamformer::ATA<CF32, CF32>> beamformer;
    std::shared_ptr<Modules::Beamformer::ATA<CI8, CF32>> beamformerCi8;
    std::shared_ptr<Modules::Polarizer<CF32, CF32>> polarizer;
    std::shared_ptr<Modules::Detector<CF32, F32>> detector;
    std::shared_ptr<Modules::Cast<CF32, OT>> complexOutputCast;
    std::shared_ptr<Modules::Cast<F32, OT>> outputCast;

    const ArrayTensor<Device::CUDA, CF32>& beamformerOutput() const {
        return beamformer ? beamformer->getOutputBuffer() : beamformerCi8->getOutputBuffer();
    }
};

}  // namespace Blade::Bundles::ATA

#endif
