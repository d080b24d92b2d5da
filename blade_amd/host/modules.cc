// Module implementations of the Blade C++ API over the gfx950 C-ABI
// (blade_amd.h).  Constructors validate like the reference and throw a
// Result; process() launches on the given stream and returns a Result.
#include "blade/base.hh"

#include <algorithm>
#include <unordered_set>

namespace Blade {
namespace {

// Streams passed to process() default to the null stream.
inline blade_stream_t abi(const Stream& s) { return s.abi(); }

}  // namespace

// ---------------------------------------------------------------- Beamformer
// (reference src/modules/beamformer/generic.cc:9-33, ata.cc:7-72)

namespace Modules::Beamformer {

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "M::BEAMFORMER"

template <typename IT, typename OT>
Generic<IT, OT>::Generic(const Config& config, const Input& input, const Stream&)
    : config(config), input(input) {
    if (config.blockSize == 0) {   // the reference divides by it unchecked (generic.cc:17)
        BL_FATAL("The block size should be more than zero.");
        throw Result::ERROR;
    }
    if (getInputBuffer().shape().numberOfTimeSamples() % config.blockSize != 0) {
        BL_FATAL("Number of time samples ({}) isn't divisable by the block size ({}).",
                 getInputBuffer().shape().numberOfTimeSamples(), config.blockSize);
        throw Result::ERROR;
    }
    if (config.enableSplitFp16 && !std::is_same<IT, CI8>::value) {
        BL_FATAL("Split-FP16 beamforming needs CI8 input (exact FP16 widening).");
        throw Result::ERROR;
    }
    BL_INFO("Type: {} -> {}", TypeInfo<IT>::name, TypeInfo<OT>::name);
    BL_INFO("Phasors Shape: {}", getInputPhasors().shape());
    BL_INFO("Enable Incoherent Beam: {}", config.enableIncoherentBeam ? "YES" : "NO");
    BL_INFO("Enable Incoherent Beam Square Root: {}", config.enableIncoherentBeamSqrt ? "YES" : "NO");
    BL_INFO("Arithmetic: {}", config.enableSplitFp16 ? "split FP16 MFMA" : "FP32 MFMA");
}

template <typename IT, typename OT>
Result Generic<IT, OT>::process(const U64&, const Stream& stream) {
    const auto& s = getInputBuffer().shape();
    const int prec = config.enableSplitFp16 ? BLADE_PREC_F16X2 : BLADE_PREC_F32;
    const U64 B = getInputPhasors().shape().numberOfBeams();
    if (workspace.empty()) {
        const U64 bytes = blade_beamformer_workspace_size(prec, s.numberOfAspects(),
                                                          s.numberOfFrequencyChannels(), B);
        workspace = Vector<Device::CUDA, U8, VectorShape>({bytes});
    }
    BL_ABI_CHECK(blade_beamform_chain_ws(
        TypeInfo<IT>::abi, prec, input.buf.data(), input.phasors.data(), output.buf.data(),
        s.numberOfAspects(), s.numberOfFrequencyChannels(), s.numberOfTimeSamples(),
        s.numberOfPolarizations(), B, config.enableIncoherentBeam, config.enableIncoherentBeamSqrt,
        epilogueRate, epiloguePolarize ? 1 : 0, workspace.data(), workspace.size(), abi(stream)));
    return Result::SUCCESS;
}

template <typename IT, typename OT>
Result Generic<IT, OT>::fuseEpilogue(U64 rate, bool polarize) {
    const U64 T = getInputBuffer().shape().numberOfTimeSamples();
    if (rate != 1 && rate != 2 && rate != 4) return Result::ERROR;
    if (T % rate != 0 || (rate > 1 && T % 4 != 0)) return Result::ERROR;
    epilogueRate = rate;
    epiloguePolarize = polarize;
    return Result::SUCCESS;
}

template <typename IT, typename OT>
ATA<IT, OT>::ATA(const typename Generic<IT, OT>::Config& config,
                 const typename Generic<IT, OT>::Input& input, const Stream& stream)
    : Generic<IT, OT>(config, input, stream) {
    const auto& ph = this->getInputPhasors().shape();
    const auto& in = this->getInputBuffer().shape();
    if (ph.numberOfBeams() > config.blockSize) {
        BL_FATAL("The block size ({}) is smaller than the number of beams ({}).", config.blockSize,
                 ph.numberOfBeams());
        throw Result::ERROR;
    }
    if (ph.numberOfFrequencyChannels() != in.numberOfFrequencyChannels()) {
        BL_FATAL("Number of frequency channels mismatch between phasors ({}) and buffer ({}).",
                 ph.numberOfFrequencyChannels(), in.numberOfFrequencyChannels());
        throw Result::ERROR;
    }
    if (ph.numberOfPolarizations() != in.numberOfPolarizations()) {
        BL_FATAL("Number of polarizations mismatch between phasors ({}) and buffer ({}).",
                 ph.numberOfPolarizations(), in.numberOfPolarizations());
        throw Result::ERROR;
    }
    if (ph.numberOfAntennas() != in.numberOfAspects()) {
        BL_FATAL("Number of antennas mismatch between phasors ({}) and buffer ({}).",
                 ph.numberOfAntennas(), in.numberOfAspects());
        throw Result::ERROR;
    }
    if (in.numberOfPolarizations() != 2) {
        BL_FATAL("Number of polarizations ({}) must be two.", in.numberOfPolarizations());
        throw Result::ERROR;
    }
    this->output.buf = ArrayTensor<Device::CUDA, OT>(this->getOutputBufferShape());
    // Allocated here so that process() allocates nothing (hipGraph capture).
    const U64 wsBytes = blade_beamformer_workspace_size(
        config.enableSplitFp16 ? BLADE_PREC_F16X2 : BLADE_PREC_F32, in.numberOfAspects(),
        in.numberOfFrequencyChannels(), ph.numberOfBeams());
    this->workspace = Vector<Device::CUDA, U8, VectorShape>({wsBytes});
    BL_INFO("Shape: {} -> {}", in, this->getOutputBuffer().shape());
}

template class BLADE_API Generic<CF32, CF32>;
template class BLADE_API Generic<CI8, CF32>;
template class BLADE_API ATA<CF32, CF32>;
template class BLADE_API ATA<CI8, CF32>;

template <typename IT, typename OT>
Batched<IT, OT>::Batched(const Config& config, const Input& input, const Stream&)
    : config(config), input(input) {
    static_assert(std::is_same<IT, CI8>::value, "the batched beamformer takes CI8 voltages");
    if (input.bufs.size() != input.phasors.size()) {
        BL_FATAL("Batched beamformer: {} input buffers but {} phasor tensors.", input.bufs.size(),
                 input.phasors.size());
        throw Result::ERROR;
    }
    for (U64 k = 0; k < input.bufs.size(); k++) {
        const auto& in = input.bufs[k].shape();
        const auto& ph = input.phasors[k].shape();
        if (ph.numberOfAntennas() != in.numberOfAspects() ||
            ph.numberOfFrequencyChannels() != in.numberOfFrequencyChannels() ||
            ph.numberOfPolarizations() != in.numberOfPolarizations() || in.numberOfPolarizations() != 2) {
            BL_FATAL("Batched beamformer: block {} has input {} and phasors {}.", k, in, ph);
            throw Result::ERROR;
        }
        output.bufs.emplace_back(ArrayShape({ph.numberOfBeams(), in.numberOfFrequencyChannels(),
                                             in.numberOfTimeSamples(), 2}));
        table.push_back(blade_beamform_block{input.bufs[k].data(), input.phasors[k].data(),
                                             output.bufs[k].data(), in.numberOfAspects(),
                                             in.numberOfFrequencyChannels(), in.numberOfTimeSamples(),
                                             ph.numberOfBeams()});
    }
    const int prec = config.enableSplitFp16 ? BLADE_PREC_F16X2 : BLADE_PREC_F32;
    const U64 bytes = blade_beamform_batched_workspace_size(prec, table.data(), table.size(), &config.tuning);
    if (bytes == 0 && !table.empty()) {
        BL_FATAL("Batched beamformer: {}", blade_amd_last_error());
        throw Result::ERROR;
    }
    workspace = Vector<Device::CUDA, U8, VectorShape>({std::max<U64>(bytes, 256)});
    BL_INFO("Blocks: {}, workspace {} bytes, arithmetic {}", table.size(), bytes,
            config.enableSplitFp16 ? "split FP16 MFMA" : "FP32 MFMA");
}

template <typename IT, typename OT>
Result Batched<IT, OT>::process(const U64&, const Stream& stream) {
    const int prec = config.enableSplitFp16 ? BLADE_PREC_F16X2 : BLADE_PREC_F32;
    BL_ABI_CHECK(blade_beamform_batched(prec, table.data(), table.size(), &config.tuning,
                                        workspace.data(), workspace.size(), abi(stream)));
    return Result::SUCCESS;
}

template class BLADE_API Batched<CI8, CF32>;

}  // namespace Modules::Beamformer

// --------------------------------------------------------------- Channelizer
// (reference src/modules/channelizer/base.cc:9-146)

namespace Modules {

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "M::CHANNELIZER"

template <typename IT, typename OT>
Channelizer<IT, OT>::Channelizer(const Config& config, const Input& input, const Stream& stream)
    : config(config), input(input), isFused(false) {
    const auto& s = getInputBuffer().shape();
    if (config.rate == 0 || s.numberOfTimeSamples() % config.rate != 0) {
        BL_FATAL("The number of time samples ({}) should be divisable by the channelizer rate ({}).",
                 s.numberOfTimeSamples(), config.rate);
        throw Result::ERROR;
    }
    if ((config.rate % 2) != 0 && config.rate != 1) {
        BL_FATAL("The channelizer rate ({}) should be divisable by 2.", config.rate);
        throw Result::ERROR;
    }
    if (s.numberOfPolarizations() != 2 && s.numberOfPolarizations() != 1) {
        BL_FATAL("Number of polarizations ({}) of the input should be one or two.",
                 s.numberOfPolarizations());
        throw Result::ERROR;
    }
    // In place for every rate (reference base.cc:45).  The spectra of one
    // (aspect, channel) occupy the same T*P samples as its input, but for
    // 1 < rate < T in another order, so the transform writes a staging
    // buffer first.
    BL_CHECK_THROW(Link(output.buf, input.buf, getOutputBufferShape()));
    if (config.rate != 1 && config.rate != s.numberOfTimeSamples()) {
        staging = ArrayTensor<Device::CUDA, OT>(getOutputBufferShape());
    }
    if (config.rate > 1 && !Profiler::IsCapturing()) {
        if (blade_channelizer_plan(config.rate, stream.abi()) != 0) {
            BL_FATAL("{}", blade_amd_last_error());
            throw Result::ERROR;
        }
    }
    BL_INFO("Type: {} -> {}", TypeInfo<IT>::name, TypeInfo<OT>::name);
    BL_INFO("Shape: {} -> {}", s, getOutputBuffer().shape());
    BL_INFO("FFT Size: {}", config.rate);
}

template <typename IT, typename OT>
Result Channelizer<IT, OT>::process(const U64&, const Stream& stream) {
    // Follow the input handle: a Runner re-points a Duet to the block's slot
    // before each cycle (the reference links once, in the constructor).
    if (output.buf.data() != input.buf.data()) BL_CHECK(Link(output.buf, input.buf, getOutputBufferShape()));
    if (config.rate == 1 || isFused) return Result::SUCCESS;
    const auto& s = getInputBuffer().shape();
    if (staging.empty()) {
        BL_ABI_CHECK(blade_channelizer(input.buf.data(), output.buf.data(), s.numberOfAspects(),
                                       s.numberOfFrequencyChannels(), s.numberOfTimeSamples(),
                                       s.numberOfPolarizations(), config.rate, abi(stream)));
        return Result::SUCCESS;
    }
    BL_ABI_CHECK(blade_channelizer(input.buf.data(), staging.data(), s.numberOfAspects(),
                                   s.numberOfFrequencyChannels(), s.numberOfTimeSamples(),
                                   s.numberOfPolarizations(), config.rate, abi(stream)));
    BL_ABI_CHECK(blade_gather(staging.data(), output.buf.data(), 1, staging.size_bytes(), 1, 0,
                              abi(stream)));
    return Result::SUCCESS;
}

template <typename IT, typename OT>
void Channelizer<IT, OT>::fuseIntoProducer() {
    isFused = true;
    staging = ArrayTensor<Device::CUDA, OT>();   // not needed any more
}

template class BLADE_API Channelizer<CF32, CF32>;

// ------------------------------------------------------- ChannelizerDetector

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "M::CHANNELIZER_DETECTOR"

template <typename IT, typename OT>
bool ChannelizerDetector<IT, OT>::Supports(U64 rate) {
    if (rate == 0 || rate > 8192) return false;
    U64 r = rate;
    for (U64 p : {2, 3, 5, 7}) {
        while (r % p == 0) r /= p;
    }
    return r == 1;
}

template <typename IT, typename OT>
ChannelizerDetector<IT, OT>::ChannelizerDetector(const Config& config, const Input& input, const Stream& stream)
    : config(config), input(input), steps(1) {
    const auto& s = getInputBuffer().shape();
    if (config.rate == 0 || s.numberOfTimeSamples() % config.rate != 0) {
        BL_FATAL("The number of time samples ({}) should be divisable by the channelizer rate ({}).",
                 s.numberOfTimeSamples(), config.rate);
        throw Result::ERROR;
    }
    if ((config.rate % 2) != 0 && config.rate != 1) {
        BL_FATAL("The channelizer rate ({}) should be divisable by 2.", config.rate);
        throw Result::ERROR;
    }
    if (!Supports(config.rate)) {
        BL_FATAL("The fused channelizer-detector does not run rate {}.", config.rate);
        throw Result::ERROR;
    }
    if (s.numberOfPolarizations() != 2) {
        BL_FATAL("Number of polarizations ({}) should be two (2).", s.numberOfPolarizations());
        throw Result::ERROR;
    }
    if (config.integrationSize == 0) {
        BL_WARN("Integration size ({}) should be more than zero.", config.integrationSize);
        throw Result::ERROR;
    }
    const U64 S = spectra();
    if (config.integrationSize > S && config.integrationSize % S == 0) {
        steps = config.integrationSize / S;
    } else if (S % config.integrationSize != 0) {
        BL_FATAL("The number of spectra ({}) should be divisable by the integration size ({}).", S,
                 config.integrationSize);
        throw Result::ERROR;
    }
    if (config.numberOfOutputPolarizations != 4 && config.numberOfOutputPolarizations != 1) {
        BL_FATAL("Number of output polarizations ({}) not supported.", config.numberOfOutputPolarizations);
        throw Result::ERROR;
    }
    output.buf = ArrayTensor<Device::CUDA, OT>(getOutputBufferShape());
    if (!Profiler::IsCapturing() && blade_channelizer_plan(config.rate, stream.abi()) != 0) {
        BL_FATAL("{}", blade_amd_last_error());
        throw Result::ERROR;
    }
    BL_INFO("Type: {} -> {}", TypeInfo<IT>::name, TypeInfo<OT>::name);
    BL_INFO("Shape: {} -> {}", s, getOutputBuffer().shape());
    BL_INFO("FFT Size: {}, Circular: {}, Integration Size: {} ({} block(s))", config.rate,
            config.convertToCircular ? "YES" : "NO", config.integrationSize, steps);
}

template <typename IT, typename OT>
Result ChannelizerDetector<IT, OT>::process(const U64& currentStepCount, const Stream& stream) {
    const auto& s = getInputBuffer().shape();
    const int accumulate = (steps > 1 && currentStepCount > 0) ? 1 : 0;
    BL_ABI_CHECK(blade_channelizer_detector(input.buf.data(), output.buf.data(), s.numberOfAspects(),
                                            s.numberOfFrequencyChannels(), s.numberOfTimeSamples(),
                                            config.rate, config.integrationSize,
                                            config.numberOfOutputPolarizations,
                                            config.convertToCircular ? 1 : 0, accumulate, abi(stream)));
    return Result::SUCCESS;
}

template class BLADE_API ChannelizerDetector<CF32, F32>;

// ----------------------------------------------------------------- Polarizer
// (reference src/modules/polarizer/base.cc:12-71)

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "M::POLARIZER"

template <typename IT, typename OT>
Polarizer<IT, OT>::Polarizer(const Config& config, const Input& input, const Stream&)
    : config(config), input(input) {
    if constexpr (!std::is_same<IT, OT>::value) {
        BL_FATAL("This module requires the type of the input ({}) and output ({}) to be the same.",
                 TypeInfo<IT>::name, TypeInfo<OT>::name);
        throw Result::ERROR;
    } else {
        if (config.inputPolarization == config.outputPolarization) BL_INFO("Bypass: Enabled");
        BL_CHECK_THROW(Link(output.buf, input.buf));
        BL_INFO("Type: {} -> {}", TypeInfo<IT>::name, TypeInfo<OT>::name);
        BL_INFO("Shape: {} -> {}", getInputBuffer().shape(), getOutputBuffer().shape());
    }
}

template <typename IT, typename OT>
Result Polarizer<IT, OT>::process(const U64&, const Stream& stream) {
    if constexpr (std::is_same<IT, OT>::value) {
        if (output.buf.data() != input.buf.data()) BL_CHECK(Link(output.buf, input.buf));   // see Channelizer
    }
    if (config.inputPolarization == config.outputPolarization || isFused) return Result::SUCCESS;
    BL_ABI_CHECK(blade_polarizer(TypeInfo<IT>::abi, input.buf.data(), output.buf.data(),
                                 input.buf.size(), abi(stream)));
    return Result::SUCCESS;
}

template class BLADE_API Polarizer<CF32, CF32>;
template class BLADE_API Polarizer<CF16, CF16>;

// ------------------------------------------------------------------ Detector
// (reference src/modules/detector/base.cc:9-87)

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "M::DETECTOR"

template <typename IT, typename OT>
Detector<IT, OT>::Detector(const Config& config, const Input& input, const Stream&)
    : config(config), input(input), steps(1) {
    const auto& s = getInputBuffer().shape();
    if (config.integrationSize <= 0) {
        BL_WARN("Integration size ({}) should be more than zero.", config.integrationSize);
        throw Result::ERROR;
    }
    const U64 T = s.numberOfTimeSamples();
    if (config.integrationSize > T && T > 0 && config.integrationSize % T == 0) {
        steps = config.integrationSize / T;   // stepped integration across blocks
    } else if (T % config.integrationSize != 0) {
        BL_FATAL("The number of time samples ({}) should be divisable by the integration size ({}).",
                 T, config.integrationSize);
        throw Result::ERROR;
    }
    if (s.numberOfPolarizations() != 2) {
        BL_FATAL("Number of polarizations ({}) should be two (2).", s.numberOfPolarizations());
        throw Result::ERROR;
    }
    if (s.numberOfAspects() <= 0) {
        BL_FATAL("Number of aspects ({}) should be more than zero.", s.numberOfAspects());
        throw Result::ERROR;
    }
    if (config.numberOfOutputPolarizations != 4 && config.numberOfOutputPolarizations != 1) {
        BL_FATAL("Number of output polarizations ({}) not supported.", config.numberOfOutputPolarizations);
        throw Result::ERROR;
    }
    output.buf = ArrayTensor<Device::CUDA, OT>(getOutputBufferShape());
    BL_INFO("Type: {} -> {}", TypeInfo<IT>::name, TypeInfo<OT>::name);
    BL_INFO("Shape: {} -> {}", s, getOutputBuffer().shape());
    BL_INFO("Integration Size: {} ({} block(s))", config.integrationSize, steps);
    BL_INFO("Number of Output Polarizations: {}", config.numberOfOutputPolarizations);
}

template <typename IT, typename OT>
Result Detector<IT, OT>::process(const U64& currentStepCount, const Stream& stream) {
    const auto& s = getInputBuffer().shape();
    const U64 T = s.numberOfTimeSamples();
    const U64 intg = (steps > 1) ? T : config.integrationSize;
    const int accumulate = (steps > 1 && currentStepCount > 0) ? 1 : 0;
    BL_ABI_CHECK(blade_detector(input.buf.data(), output.buf.data(), s.numberOfAspects(),
                                s.numberOfFrequencyChannels(), T, intg,
                                config.numberOfOutputPolarizations, accumulate, abi(stream)));
    return Result::SUCCESS;
}

template class BLADE_API Detector<CF32, F32>;

// ---------------------------------------------------------------------- Cast
// (reference src/modules/cast/base.cc:12-100)

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "M::CAST"

template <typename IT, typename OT>
Cast<IT, OT>::Cast(const Config& config, const Input& input, const Stream&)
    : config(config), input(input) {
    if (TypeInfo<IT>::is_complex != TypeInfo<OT>::is_complex) {
        BL_FATAL("Cannot cast between complex and non-complex types.");
        throw Result::ERROR;
    }
    if constexpr (std::is_same<IT, OT>::value) {
        BL_DEBUG("Bypassing cast because input and output types are the same.");
        BL_CHECK_THROW(Link(output.buf, input.buf));
    } else {
        output.buf = ArrayTensor<Device::CUDA, OT>(getInputBuffer().shape());
        BL_INFO("Type: {} -> {}", TypeInfo<IT>::name, TypeInfo<OT>::name);
        BL_INFO("Shape: {} -> {}", getInputBuffer().shape(), getOutputBuffer().shape());
    }
}

template <typename IT, typename OT>
Result Cast<IT, OT>::process(const U64&, const Stream& stream) {
    if constexpr (std::is_same<IT, OT>::value) {
        if (output.buf.data() != input.buf.data()) BL_CHECK(Link(output.buf, input.buf));   // see Channelizer
        return Result::SUCCESS;
    } else {
        BL_ABI_CHECK(blade_cast(TypeInfo<IT>::abi, TypeInfo<OT>::abi, input.buf.data(),
                                output.buf.data(), input.buf.size(), abi(stream)));
        return Result::SUCCESS;
    }
}

template class BLADE_API Cast<I8, I8>;
template class BLADE_API Cast<I8, F32>;
template class BLADE_API Cast<F16, F16>;
template class BLADE_API Cast<F16, F32>;
template class BLADE_API Cast<F16, CF32>;
template class BLADE_API Cast<F16, CF16>;
template class BLADE_API Cast<F32, I32>;
template class BLADE_API Cast<F32, F32>;
template class BLADE_API Cast<F32, F16>;
template class BLADE_API Cast<F32, CF32>;
template class BLADE_API Cast<F32, CF16>;
template class BLADE_API Cast<CI8, CI8>;
template class BLADE_API Cast<CI8, CF32>;
template class BLADE_API Cast<CI8, CF16>;
template class BLADE_API Cast<CF32, F16>;
template class BLADE_API Cast<CF32, F32>;
template class BLADE_API Cast<CF32, CF16>;
template class BLADE_API Cast<CF32, CF32>;
template class BLADE_API Cast<CF16, F16>;
template class BLADE_API Cast<CF16, F32>;
template class BLADE_API Cast<CF16, CF32>;
template class BLADE_API Cast<CF16, CF16>;

// ----------------------------------------------------------------- Duplicate
// (reference src/modules/duplicate/base.cc:12-42)

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "M::DUPLICATE"

template <typename IT, typename OT>
Duplicate<IT, OT>::Duplicate(const Config& config, const Input& input, const Stream&)
    : config(config), input(input) {
    if constexpr (!std::is_same<IT, OT>::value) {
        BL_FATAL("Input ({}) and output ({}) types aren't the same. Casting isn't supported by Duplicate yet.",
                 TypeInfo<IT>::name, TypeInfo<OT>::name);
        throw Result::ERROR;
    }
    output.buf = ArrayTensor<Device::CUDA, OT>(input.buf.shape());
    BL_INFO("Type: {} -> {}", TypeInfo<IT>::name, TypeInfo<OT>::name);
    BL_INFO("Shape: {} -> {}", getInputBuffer().shape(), getOutputBuffer().shape());
}

template <typename IT, typename OT>
Result Duplicate<IT, OT>::process(const U64&, const Stream& stream) {
    // One row of all the bytes, one step: the Gather copy kernel as a plain copy.
    BL_ABI_CHECK(blade_gather(input.buf.data(), output.buf.data(), 1, input.buf.size() * sizeof(IT), 1, 0,
                              abi(stream)));
    return Result::SUCCESS;
}

template class BLADE_API Duplicate<CI8, CI8>;
template class BLADE_API Duplicate<CF16, CF16>;
template class BLADE_API Duplicate<CF32, CF32>;
template class BLADE_API Duplicate<F32, F32>;

// -------------------------------------------------------------------- Gather
// (reference src/modules/gather/base.cc:12-116)

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "M::GATHER"

template <typename IT, typename OT>
Gather<IT, OT>::Gather(const Config& config, const Input& input, const Stream&)
    : config(config), input(input), computeRatio(config.multiplier) {
    if constexpr (!std::is_same<IT, OT>::value) {
        BL_FATAL("Input ({}) and output ({}) types aren't the same. Casting isn't supported by Gather yet.",
                 TypeInfo<IT>::name, TypeInfo<OT>::name);
        throw Result::ERROR;
    }
    if (config.axis >= input.buf.shape().dimensions()) {
        BL_FATAL("Selected input axis ({}) is larger than input shape dimensions ({}).", config.axis,
                 input.buf.shape().dimensions());
        throw Result::ERROR;
    }
    if (config.multiplier <= 0) {
        BL_FATAL("Multiplier ({}) should be more than zero.", config.multiplier);
        throw Result::ERROR;
    }
    widthSize = 1;
    for (U64 i = config.axis; i < input.buf.shape().dimensions(); i++) widthSize *= input.buf.shape()[i];
    widthByteSize = widthSize * sizeof(IT);
    heightSize = 1;
    for (U64 i = 0; i < config.axis; i++) heightSize *= input.buf.shape()[i];
    output.buf = ArrayTensor<Device::CUDA, OT>(getOutputBufferShape());
    BL_INFO("Type: {} -> {}", TypeInfo<IT>::name, TypeInfo<OT>::name);
    BL_INFO("Shape: {} -> {}", getInputBuffer().shape(), getOutputBuffer().shape());
    BL_INFO("Axis: {}", config.axis);
    BL_INFO("Multiplier: {}", computeRatio);
}

template <typename IT, typename OT>
Result Gather<IT, OT>::process(const U64& currentStepCount, const Stream& stream) {
    BL_ABI_CHECK(blade_gather(input.buf.data(), output.buf.data(), heightSize, widthByteSize,
                              computeRatio, currentStepCount, abi(stream)));
    return Result::SUCCESS;
}

template class BLADE_API Gather<CI8, CI8>;
template class BLADE_API Gather<CF16, CF16>;
template class BLADE_API Gather<CF32, CF32>;
template class BLADE_API Gather<F32, F32>;

// --------------------------------------------------------------- Permutation
// (reference src/modules/permutation/base.cc:12-77)

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "M::PERMUTATION"

template <typename IT, typename OT>
Permutation<IT, OT>::Permutation(const Config& config, const Input& input, const Stream&)
    : config(config), input(input) {
    if constexpr (!std::is_same<IT, OT>::value) {
        BL_FATAL("Input ({}) and output ({}) types aren't the same. Casting isn't supported by Permutation yet.",
                 TypeInfo<IT>::name, TypeInfo<OT>::name);
        throw Result::ERROR;
    }
    std::unordered_set<U64> seen;
    for (U64 i = 0; i < config.indexes.dimensions(); i++) {
        if (config.indexes[i] >= 4) {
            BL_FATAL("Permutation index {} is out of range.", config.indexes[i]);
            throw Result::ERROR;
        }
        if (seen.count(config.indexes[i])) {
            BL_FATAL("Repeated index {} in permutation.", config.indexes[i]);
            throw Result::ERROR;
        }
        seen.insert(config.indexes[i]);
    }
    output.buf = ArrayTensor<Device::CUDA, OT>(getOutputBufferShape());
    BL_INFO("Type: {} -> {}", TypeInfo<IT>::name, TypeInfo<OT>::name);
    BL_INFO("Shape: {} -> {}", getInputBuffer().shape(), getOutputBuffer().shape());
    BL_INFO("Indexes: {}", config.indexes);
}

template <typename IT, typename OT>
Result Permutation<IT, OT>::process(const U64&, const Stream& stream) {
    const auto& s = getInputBuffer().shape();
    const uint64_t shape[4] = {s[0], s[1], s[2], s[3]};
    const uint64_t idx[4] = {config.indexes[0], config.indexes[1], config.indexes[2], config.indexes[3]};
    BL_ABI_CHECK(blade_permutation(input.buf.data(), output.buf.data(), sizeof(IT), shape, idx,
                                   abi(stream)));
    return Result::SUCCESS;
}

template class BLADE_API Permutation<CI8, CI8>;
template class BLADE_API Permutation<CF16, CF16>;
template class BLADE_API Permutation<CF32, CF32>;

}  // namespace Modules
}  // namespace Blade
