// Generic Mode-H bundle (reference include/blade/bundles/generic/mode_h.hh:13-126):
// Cast -> Channelizer(rate = T) -> Polarizer -> Detector -> Cast.  The middle
// three run as one fused ChannelizerDetector launch whenever its kernel
// supports the rate (T <= 8192 made of the factors 2, 3, 5, 7): their
// buffers are private to the bundle, so only the launch count changes.
// (The reference's in-place chain also overwrites a CF32 input buffer with
// its spectra; the fused launch leaves the input as it was.)
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
    bool fused() const { return channelizerDetector != nullptr; }

    explicit ModeH(const Config& config, const Input& input, const Stream& stream)
        : Bundle(stream), config(config), input(input) {
        const U64 rate = config.inputShape.numberOfTimeSamples();
        this->connect(inputCast, {.blockSize = config.castBlockSize}, {.buf = this->input.buffer});
        if (Modules::ChannelizerDetector<CF32, F32>::Supports(rate)) {
            this->connect(channelizerDetector,
                          {.rate = rate,
                           .convertToCircular = config.polarizerConvertToCircular,
                           .integrationSize = config.detectorIntegrationSize,
                           .numberOfOutputPolarizations = config.detectorNumberOfOutputPolarizations,
                           .blockSize = config.detectorBlockSize},
                          {.buf = inputCast->getOutputBuffer()});
            this->connect(outputCast, {.blockSize = config.castBlockSize},
                          {.buf = channelizerDetector->getOutputBuffer()});
        } else {
            this->connect(channelizer, {.rate = rate, .blockSize = config.channelizerBlockSize},
                          {.buf = inputCast->getOutputBuffer()});
            this->connect(polarizer, {.inputPolarization = POL::XY,
                                      .outputPolarization = config.polarizerConvertToCircular ? POL::LR : POL::XY,
                                      .blockSize = config.polarizerBlockSize},
                          {.buf = channelizer->getOutputBuffer()});
            this->connect(detector, {.integrationSize = config.detectorIntegrationSize,
                                     .numberOfOutputPolarizations = config.detectorNumberOfOutputPolarizations,
                                     .blockSize = config.detectorBlockSize},
                          {.buf = polarizer->getOutputBuffer()});
            this->connect(outputCast, {.blockSize = config.castBlockSize},
                          {.buf = detector->getOutputBuffer()});
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
    std::shared_ptr<Modules::Polarizer<CF32, CF32>> polarizer;
    std::shared_ptr<Modules::Detector<CF32, F32>> detector;
    std::shared_ptr<Modules::ChannelizerDetector<CF32, F32>> channelizerDetector;
    std::shared_ptr<Modules::Cast<F32, OT>> outputCast;
};

}  // namespace Blade::Bundles::Generic

#endif
