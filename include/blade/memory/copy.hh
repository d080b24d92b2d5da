// Stream-ordered copies between tensors (reference include/blade/memory/copy.hh:11-181).
// The HIP runtime infers the direction from the pointers (hipMemcpyDefault),
// so every Copy/Copy2D overload uses the right kind (the reference passes
// DeviceToHost for every 2-D copy, copy.hh:151-180).
//
// Host-to-host copies have finished when Copy returns, as CUDA's
// cudaMemcpyAsync does for them (the reference's CPU -> CPU Copy relies on
// it: modules read host inputs such as the Julian date in process(), right
// after the Runner's input callback copied them).  HIP would queue them on
// the stream instead, so here the stream's earlier work is awaited and the
// bytes are copied by the host.
#ifndef BLADE_MEMORY_COPY_HH
#define BLADE_MEMORY_COPY_HH

#include <hip/hip_runtime_api.h>

#include <cstring>

#include "blade/macros.hh"
#include "blade/memory/vector.hh"

namespace Blade {

template <Device D, Device S, typename Type, typename Dims>
inline Result Copy(Vector<D, Type, Dims>& dst, const Vector<S, Type, Dims>& src,
                   const Stream& stream = {}) {
    if (dst.size() != src.size()) {
        BL_FATAL("Size mismatch between source and destination ({}, {}).", src.size(), dst.size());
        return Result::ASSERTION_ERROR;
    }
    if (dst.shape() != src.shape()) {
        BL_FATAL("Shape mismatch between source ({}) and destination ({}).", src.shape(), dst.shape());
        return Result::ASSERTION_ERROR;
    }
    if (src.size_bytes() == 0 || dst.data() == src.data()) return Result::SUCCESS;
    if constexpr (D == Device::CPU && S == Device::CPU) {
        if (!dst.unified() && !src.unified()) {
            BL_HIP_CHECK(hipStreamSynchronize(stream.handle));
            std::memcpy(dst.data(), src.data(), src.size_bytes());
            return Result::SUCCESS;
        }
    }
    BL_HIP_CHECK(hipMemcpyAsync(dst.data(), src.data(), src.size_bytes(), hipMemcpyDefault,
                                stream.handle));
    return Result::SUCCESS;
}

// Strided 2-D copy: `height` rows of `width` bytes; pitches and byte offsets
// on each side.
template <Device D, Device S, typename Type, typename Dims>
inline Result Copy2D(Vector<D, Type, Dims>& dst, U64 dstPitch, U64 dstPad,
                     const Vector<S, Type, Dims>& src, U64 srcPitch, U64 srcPad,
                     U64 width, U64 height, const Stream& stream = {}) {
    if (width > dstPitch || width > srcPitch) {
        BL_FATAL("2-D copy width ({}) exceeds a pitch ({}, {}).", width, dstPitch, srcPitch);
        return Result::ASSERTION_ERROR;
    }
    if (width == 0 || height == 0) return Result::SUCCESS;
    auto* d = reinterpret_cast<uint8_t*>(dst.data()) + dstPad;
    const auto* s = reinterpret_cast<const uint8_t*>(src.data()) + srcPad;
    BL_HIP_CHECK(hipMemcpy2DAsync(d, dstPitch, s, srcPitch, width, height, hipMemcpyDefault,
                                  stream.handle));
    return Result::SUCCESS;
}

}  // namespace Blade

#endif
