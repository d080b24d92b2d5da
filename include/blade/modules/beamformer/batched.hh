// Batched beamformer module: several ATA beamformer blocks of different
// shapes (antennas, channels, samples, beams) in one persistent launch.  No
// reference counterpart: each block is what Modules::Beamformer::ATA computes
// (reference src/modules/beamformer/beamformer.cu:13-79) without the
// incoherent beam.  process() allocates nothing, so a pipeline holding it can
// be captured in a hipGraph.
#ifndef BLADE_MODULES_BEAMFORMER_BATCHED_HH
#define BLADE_MODULES_BEAMFORMER_BATCHED_HH

#include <string>
#include <vector>

#include "blade/module.hh"

namespace Blade::Modules::Beamformer {

template <typename IT, typename OT>
class BLADE_API Batched : public Module {
 public:
    struct Config {
        // FP32 matrix cores, or split FP16 (CI8 input).
        BOOL enableSplitFp16 = false;
        // Tile and launch shape; zeros pick them automatically.
        blade_batch_tuning tuning = {0, 0, 0, 0};
    };

    const Config& getConfig() const { return config; }

    // Tensors are reference-counted handles: the module shares them.
    struct Input {
        std::vector<ArrayTensor<Device::CUDA, IT>> bufs;
        std::vector<PhasorTensor<Device::CUDA, CF32>> phasors;
    };

    struct Output {
        std::vector<ArrayTensor<Device::CUDA, OT>> bufs;
    };

    const std::vector<ArrayTensor<Device::CUDA, OT>>& getOutputBuffers() const { return output.bufs; }

    Taint getTaint() const override { return Taint::CONSUMER | Taint::PRODUCER; }
    std::string name() const override { return "BatchedBeamformer"; }

    explicit Batched(const Config& config, const Input& input, const Stream& stream = {});
    virtual ~Batched() = default;
    Result process(const U64& currentStepCount, const Stream& stream = {}) override;

 protected:
    const Config config;
    const Input input;
    Output output;
    std::vector<blade_beamform_block> table;
    Vector<Device::CUDA, U8, VectorShape> workspace;
};

}  // namespace Blade::Modules::Beamformer

#endif
