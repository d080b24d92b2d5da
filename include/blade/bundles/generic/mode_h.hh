// Generic Mode-H bundle (reference include/blade/bundles/generic/mode_h.hh:13-126):
// Cast -> Channelizer(rate = T) -> Polarizer -> Detector -> Cast.
#ifndef BLADE_BUNDLES_GENERIC_MODE_H_HH
#define BLADE_BUNDLES_GENERIC_MODE_H_HH

#include <memory>

#include "blade/bundle.hh"
#include "blade/modules/base.hh"

namespace Blade::Bundles::Generic {

template <typename IT, typename OT>
class BLADE_API ModeH : public Bundle {
 public:
    struct Config {
        ArrayShape inputShape;
        ArrayShape outputShape;

        BOOL polarizerConvertToCircular = false;

        U64 detectorIntegrationSize;
        U64 detectorNumberOfOutputPolarizations;

        U64 castBlockSize = 512;
        U64 polarizerBlockSize = 512;
        U64 channelizerBlockSize = 512;
        U64 detectorBlockSize = 512;
    };

    const Config& getConfig() const { return config; }

    struct Input {
        ArrayTensor<Device::CUDA, IT> buffer;
    };

    const ArrayTensor<Device::CUDA, IT>& getInputBuffer() const { return input.buffer; }
    const ArrayTensor<Device::CUDA, OT>& getOutputBuffer() { return outputCast->getOutputBuffer(); }

    explicit ModeH(const Config& config, const Input& input, const Stream& stream)
        : Bundle(stream), config(config), input(input) {
        this->connect(inputCast, {.blockSize = config.castBlockSize}, {.buf = this->input.buffer});
        this->connect(channelizer, {.rate = config.inputShape.numberOfTimeSamples(),
                                    .blockSize = config.channelizerBlockSize},
                      {.buf = inputCast->getOutputBuffer()});
        this->connect(polarizer, {.inputPolarization = POL::XY,
                                  .outputPolarization = config.polarizerConvertToCircular ? POL::LR : POL::XY,
                                  .blockSize = config.polarizerBlockSize},
                      {.buf = channelizer->getOutputBuffer()});
        this->connect(detector, {.integrationSize = config.detectorIntegrationSize,
                                 .numberOfOutputPolarizations = config.detectorNumberOfOutputPolarizations,
                                 .blockSize = config.detectorBlockSize},
                      {.buf = polarizer->getOutputBuffer()});
        this->connect(outputCast, {.blockSize = config.castBlockSize}, {.buf = detector->getOutputBuffer()});
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
    std::shared_ptr<Modules::Polarizer<CF32, CF32>> polarizer;
    std::shared_ptr<Modules::Detector<CF32, F32>> detector;
    std::shared_ptr<Modules::Cast<F32, OT>> outputCast;
};

}  // namespace Blade::Bundles::Generic

#endif
