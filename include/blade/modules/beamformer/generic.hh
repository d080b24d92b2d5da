// Beamformer module interface (reference include/blade/modules/beamformer/generic.hh:14-79).
#ifndef BLADE_MODULES_BEAMFORMER_GENERIC_HH
#define BLADE_MODULES_BEAMFORMER_GENERIC_HH

#include <string>

#include "blade/module.hh"

namespace Blade::Modules::Beamformer {

template <typename IT, typename OT>
class BLADE_API Generic : public Module {
 public:
    struct Config {
        BOOL enableIncoherentBeam = false;
        BOOL enableIncoherentBeamSqrt = false;
        U64 blockSize = 512;
        // Arithmetic of the contraction: FP32 matrix cores, or split FP16
        // (hi + lo phasors, FP32 accumulate; CI8 input only).
        BOOL enableSplitFp16 = false;
    };

    const Config& getConfig() const { return config; }

    struct Input {
        const ArrayTensor<Device::CUDA, IT>& buf;
        const PhasorTensor<Device::CUDA, CF32>& phasors;
    };

    const ArrayTensor<Device::CUDA, IT>& getInputBuffer() const { return input.buf; }
    const PhasorTensor<Device::CUDA, CF32>& getInputPhasors() const { return input.phasors; }

    struct Output {
        ArrayTensor<Device::CUDA, OT> buf;
    };

    const ArrayTensor<Device::CUDA, OT>& getOutputBuffer() const { return output.buf; }

    Taint getTaint() const override { return Taint::CONSUMER | Taint::PRODUCER; }
    std::string name() const override { return "Beamformer"; }

    explicit Generic(const Config& config, const Input& input, const Stream& stream = {});
    virtual ~Generic() = default;
    Result process(const U64& currentStepCount, const Stream& stream = {}) override;

    // Pipeline fusion (no reference counterpart): run the Channelizer of
    // `rate` (1, 2 or 4) and, with `polarize`, the XY -> LR Polarizer that
    // follow this module in place on its output buffer, inside this
    // module's own launch.  The buffer then ends up as it would after the
    // three modules ran one by one.
    Result fuseEpilogue(U64 rate, bool polarize);
    U64 fusedRate() const { return epilogueRate; }
    bool fusedPolarize() const { return epiloguePolarize; }

 protected:
    const Config config;
    const Input input;
    Output output;
    // Sink for the stores of idle lanes (allocated once; see the C-ABI).
    Vector<Device::CUDA, U8, VectorShape> workspace;
    U64 epilogueRate = 1;
    bool epiloguePolarize = false;

    virtual const ArrayShape getOutputBufferShape() const = 0;
};

}  // namespace Blade::Modules::Beamformer

#endif
