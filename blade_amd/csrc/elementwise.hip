// Bandwidth-bound data-movement kernels of the ATA path: Cast, Polarizer,
// Gather and Permutation.  All are bit-exact (or single-rounding casts) and
// stream HBM with 16-byte-per-lane accesses where alignment allows.
#include "common.hh"

#include <type_traits>

namespace bl {

// ---------------------------------------------------------------------------
// Cast — component-wise static_cast (reference src/modules/cast/cast.cu:5-12).
// ---------------------------------------------------------------------------

template <typename OT, typename IT>
__device__ __forceinline__ OT convert(IT v) {
    if constexpr (std::is_same<OT, _Float16>::value && std::is_same<IT, double>::value) {
        return static_cast<_Float16>(static_cast<float>(v));   // rejected on the host
    } else {
        return static_cast<OT>(v);
    }
}

template <typename IT, typename OT, int V>
__global__ __launch_bounds__(256) void cast_vec_kernel(const IT* __restrict__ in,
                                                       OT* __restrict__ out, uint64_t nvec) {
    struct alignas(sizeof(IT) * V) InV { IT v[V]; };
    struct alignas(sizeof(OT) * V) OutV { OT v[V]; };
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride) {
        InV a = reinterpret_cast<const InV*>(in)[i];
        OutV b;
#pragma unroll
        for (int k = 0; k < V; k++) b.v[k] = convert<OT>(a.v[k]);
        if constexpr (sizeof(OutV) == 16) {
            // Whole-line output nothing reads back: nontemporal stores.
            typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
            u32x4 t;
            __builtin_memcpy(&t, &b, 16);
            __builtin_nontemporal_store(t, reinterpret_cast<u32x4*>(out) + i);
        } else {
            reinterpret_cast<OutV*>(out)[i] = b;
        }
    }
}

template <typename IT, typename OT>
__global__ __launch_bounds__(256) void cast_scalar_kernel(const IT* __restrict__ in,
                                                          OT* __restrict__ out, uint64_t begin,
                                                          uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = begin + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = convert<OT>(in[i]);
}

static unsigned grid_for(uint64_t work, unsigned block) {
    // One workgroup per `block` items: the launch then sweeps memory as one
    // compact front, which streams at ~6.1 TB/s, against ~5.3 TB/s for a
    // grid-stride loop over a grid capped at 64 workgroups per CU.  The
    // kernels keep their grid-stride loops for sizes past the cap.
    uint64_t g = ceil_div(work, block);
    const uint64_t cap = 0x7fffffffull / block;
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (unsigned)g;
}

template <typename IT, typename OT>
static int launch_cast(const void* in, void* out, uint64_t n_scalars, hipStream_t s) {
    // Widening casts: 16 output bytes per lane, so that every store
    // instruction writes 64 lanes x 16 contiguous bytes.  Others: 8 scalars.
    constexpr int V = sizeof(OT) > sizeof(IT) ? (16 / sizeof(OT) > 0 ? 16 / sizeof(OT) : 1) : 8;
    const IT* src = static_cast<const IT*>(in);
    OT* dst = static_cast<OT*>(out);
    const bool aligned = ((uintptr_t)src % (sizeof(IT) * V) == 0) &&
                         ((uintptr_t)dst % (sizeof(OT) * V) == 0);
    uint64_t done = 0;
    if (aligned && n_scalars >= (uint64_t)V) {
        const uint64_t nvec = n_scalars / V;
        hipLaunchKernelGGL((cast_vec_kernel<IT, OT, V>), dim3(grid_for(nvec, 256)), dim3(256), 0,
                           s, src, dst, nvec);
        done = nvec * V;
    }
    if (done < n_scalars) {
        hipLaunchKernelGGL((cast_scalar_kernel<IT, OT>), dim3(grid_for(n_scalars - done, 256)),
                           dim3(256), 0, s, src, dst, done, n_scalars);
    }
    return check_launch("cast");
}

template <typename IT>
static int dispatch_cast_out(int out_scalar, const void* in, void* out, uint64_t n, hipStream_t s) {
    switch (out_scalar) {
        case BLADE_I8: return launch_cast<IT, int8_t>(in, out, n, s);
        case BLADE_F16: return launch_cast<IT, _Float16>(in, out, n, s);
        case BLADE_F32: return launch_cast<IT, float>(in, out, n, s);
        case BLADE_F64: return launch_cast<IT, double>(in, out, n, s);
        case BLADE_I32: return launch_cast<IT, int32_t>(in, out, n, s);
    }
    return fail("cast: unsupported output type");
}

static int scalar_of(int t) {
    switch (t) {
        case BLADE_CI8: return BLADE_I8;
        case BLADE_CF16: return BLADE_F16;
        case BLADE_CF32: return BLADE_F32;
        case BLADE_CF64: return BLADE_F64;
        default: return t;
    }
}

static bool valid_type(int t) { return t >= BLADE_I8 && t <= BLADE_CF64; }

// ---------------------------------------------------------------------------
// Polarizer XY -> LR (reference src/modules/polarizer/polarizer.cu:5-25).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void polarizer_cf32_kernel(const float4* __restrict__ in,
                                                             float4* __restrict__ out,
                                                             uint64_t pairs) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < pairs; i += stride) {
        const float4 v = in[i];   // (Xr, Xi, Yr, Yi)
        // X + jY = (Xr - Yi, Xi + Yr) ;  X - jY = (Xr + Yi, Xi - Yr)
        // Nontemporal: each wave writes whole lines and nothing reads them back.
        float* o = reinterpret_cast<float*>(out + i);
        __builtin_nontemporal_store(v.x - v.w, o);
        __builtin_nontemporal_store(v.y + v.z, o + 1);
        __builtin_nontemporal_store(v.x + v.w, o + 2);
        __builtin_nontemporal_store(v.y - v.z, o + 3);
    }
}

__global__ __launch_bounds__(256) void polarizer_cf16_kernel(const halfx8* __restrict__ in,
                                                             halfx8* __restrict__ out,
                                                             uint64_t quads) {
    // Two (X, Y) pairs per lane; arithmetic in native FP16 (one rounding per
    // component, like the reference's half-precision complex ops).
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < quads; i += stride) {
        const halfx8 v = in[i];
        halfx8 o;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const _Float16 xr = v[4 * q + 0], xi = v[4 * q + 1], yr = v[4 * q + 2], yi = v[4 * q + 3];
            o[4 * q + 0] = xr - yi;
            o[4 * q + 1] = xi + yr;
            o[4 * q + 2] = xr + yi;
            o[4 * q + 3] = xi - yr;
        }
        out[i] = o;
    }
}

__global__ void polarizer_cf16_tail_kernel(const _Float16* __restrict__ in,
                                           _Float16* __restrict__ out, uint64_t pair) {
    const _Float16 xr = in[4 * pair + 0], xi = in[4 * pair + 1], yr = in[4 * pair + 2],
                   yi = in[4 * pair + 3];
    out[4 * pair + 0] = xr - yi;
    out[4 * pair + 1] = xi + yr;
    out[4 * pair + 2] = xr + yi;
    out[4 * pair + 3] = xi - yr;
}

// ---------------------------------------------------------------------------
// Gather — strided 2-D block placement (reference src/modules/gather/base.cc:184-225).
// ---------------------------------------------------------------------------

template <typename V>
__global__ __launch_bounds__(256) void gather_kernel(const V* __restrict__ in, V* __restrict__ out,
                                                     uint64_t height, uint64_t wvec,
                                                     uint64_t out_pitch_vec, uint64_t out_off_vec) {
    const uint64_t total = height * wvec;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const uint64_t row = i / wvec;
        const uint64_t col = i - row * wvec;
        out[row * out_pitch_vec + out_off_vec + col] = in[i];
    }
}

template <typename V>
static int launch_gather(const void* in, void* out, uint64_t height, uint64_t width_bytes,
                         uint64_t multiplier, uint64_t step, hipStream_t s) {
    const uint64_t wvec = width_bytes / sizeof(V);
    hipLaunchKernelGGL((gather_kernel<V>), dim3(grid_for(height * wvec, 256)), dim3(256), 0, s,
                       static_cast<const V*>(in), static_cast<V*>(out), height, wvec,
                       wvec * multiplier, wvec * step);
    return check_launch("gather");
}

// ---------------------------------------------------------------------------
// Permutation — out = transpose(in, indexes) (reference permutation.cu:6-20).
// Each lane owns one output element (coalesced writes) and gathers its input.
// When the innermost axis is kept, whole rows move and reads are coalesced too.
// ---------------------------------------------------------------------------

struct Perm4 {
    uint64_t out_shape[4];
    uint64_t in_stride_of_out_dim[4];
};

template <typename V>
__global__ __launch_bounds__(256) void permutation_kernel(const V* __restrict__ in,
                                                          V* __restrict__ out, Perm4 p,
                                                          uint64_t total) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        uint64_t rem = i;
        const uint64_t o3 = rem % p.out_shape[3]; rem /= p.out_shape[3];
        const uint64_t o2 = rem % p.out_shape[2]; rem /= p.out_shape[2];
        const uint64_t o1 = rem % p.out_shape[1]; rem /= p.out_shape[1];
        const uint64_t o0 = rem;
        const uint64_t src = o0 * p.in_stride_of_out_dim[0] + o1 * p.in_stride_of_out_dim[1] +
                             o2 * p.in_stride_of_out_dim[2] + o3 * p.in_stride_of_out_dim[3];
        out[i] = in[src];
    }
}

// Tiled transpose of the two innermost moving axes through LDS, for
// permutations that move the innermost axis (e.g. AFTP -> ATPF).  The tensor
// is viewed as out[o0][o1][o2][o3]; `r` is the output axis whose input stride
// is 1 (the input's innermost axis) and the tile is (o_r x o3).
template <typename V>
__global__ __launch_bounds__(256) void permutation_tiled_kernel(const V* __restrict__ in,
                                                                V* __restrict__ out, Perm4 p,
                                                                int r, uint64_t outer_count) {
    constexpr int TILE = 32;
    __shared__ V tile[TILE][TILE + 1];
    // Output strides.
    const uint64_t os3 = 1, os2 = p.out_shape[3], os1 = os2 * p.out_shape[2],
                   os0 = os1 * p.out_shape[1];
    const uint64_t ostr[4] = {os0, os1, os2, os3};
    const uint64_t nr = p.out_shape[r], n3 = p.out_shape[3];
    const uint64_t tiles_r = (nr + TILE - 1) / TILE, tiles_3 = (n3 + TILE - 1) / TILE;
    const uint64_t tiles_per_outer = tiles_r * tiles_3;
    const uint64_t total_tiles = tiles_per_outer * outer_count;
    const int tx = threadIdx.x % TILE, ty = threadIdx.x / TILE;   // 32 x 8
    for (uint64_t t = blockIdx.x; t < total_tiles; t += gridDim.x) {
        uint64_t outer = t / tiles_per_outer;
        const uint64_t tt = t - outer * tiles_per_outer;
        const uint64_t br = (tt / tiles_3) * TILE, b3 = (tt % tiles_3) * TILE;
        // Decompose `outer` over the two output axes that are neither r nor 3.
        uint64_t oc[4] = {0, 0, 0, 0};
        for (int d = 2; d >= 0; d--) {
            if (d == r) continue;
            oc[d] = outer % p.out_shape[d];
            outer /= p.out_shape[d];
        }
        uint64_t in_base = 0, out_base = 0;
        for (int d = 0; d < 3; d++) {
            if (d == r) continue;
            in_base += oc[d] * p.in_stride_of_out_dim[d];
            out_base += oc[d] * ostr[d];
        }
        // Load: consecutive lanes walk the input's innermost axis (o_r).
        for (int k = ty; k < TILE; k += 8) {
            const uint64_t i3 = b3 + k, ir = br + tx;
            if (i3 < n3 && ir < nr)
                tile[k][tx] = in[in_base + ir + i3 * p.in_stride_of_out_dim[3]];
        }
        __syncthreads();
        // Store: consecutive lanes walk the output's innermost axis (o3).
        for (int k = ty; k < TILE; k += 8) {
            const uint64_t ir = br + k, i3 = b3 + tx;
            if (i3 < n3 && ir < nr) {
                // Rows of 32 elements of 8 or 16 bytes are whole lines that
                // nothing reads back: nontemporal stores (10% faster at
                // AFTP -> ATPF on [16,192,8192,2] CF32).
                V* const o = out + out_base + ir * ostr[r] + i3;
                if constexpr (sizeof(V) == 8) {
                    __builtin_nontemporal_store(*reinterpret_cast<const unsigned long long*>(&tile[tx][k]),
                                                reinterpret_cast<unsigned long long*>(o));
                } else if constexpr (sizeof(V) == 16) {
                    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
                    u32x4 t;
                    __builtin_memcpy(&t, &tile[tx][k], 16);
                    __builtin_nontemporal_store(t, reinterpret_cast<u32x4*>(o));
                } else {
                    *o = tile[tx][k];
                }
            }
        }
        __syncthreads();
    }
}

template <typename V>
static int launch_permutation(const void* in, void* out, const Perm4& p, hipStream_t s) {
    const uint64_t total = p.out_shape[0] * p.out_shape[1] * p.out_shape[2] * p.out_shape[3];
    int r = -1;
    for (int d = 0; d < 3; d++)
        if (p.in_stride_of_out_dim[d] == 1 && p.out_shape[d] > 1) r = d;
    // The tiled transpose when both tile axes fill most of a 32 x 32 tile;
    // otherwise the direct gather (coalesced stores, loads through L2).
    if (r >= 0 && p.in_stride_of_out_dim[3] != 1 && p.out_shape[3] >= 16 && p.out_shape[r] >= 16) {
        uint64_t outer = 1;
        for (int d = 0; d < 3; d++)
            if (d != r) outer *= p.out_shape[d];
        const uint64_t tiles = outer * ceil_div(p.out_shape[r], 32) * ceil_div(p.out_shape[3], 32);
        hipLaunchKernelGGL((permutation_tiled_kernel<V>), dim3(grid_for(tiles, 1)), dim3(256), 0, s,
                           static_cast<const V*>(in), static_cast<V*>(out), p, r, outer);
    } else {
        hipLaunchKernelGGL((permutation_kernel<V>), dim3(grid_for(total, 256)), dim3(256), 0, s,
                           static_cast<const V*>(in), static_cast<V*>(out), p, total);
    }
    return check_launch("permutation");
}

struct alignas(16) Bytes16 { uint32_t w[4]; };

}  // namespace bl

using namespace bl;

extern "C" int blade_cast(int in_type, int out_type, const void* in, void* out,
                          uint64_t n_elements, blade_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!valid_type(in_type) || !valid_type(out_type)) return fail("cast: invalid type");
    if (type_components(in_type) != type_components(out_type))
        return fail("Cannot cast between complex and non-complex types.");
    if ((scalar_of(out_type) == BLADE_F16) && (scalar_of(in_type) == BLADE_F64))
        return fail("cast: F64 -> F16 is not supported (no single-rounding conversion)");
    if (n_elements == 0) return BLADE_SUCCESS;
    if (!in || !out) return fail("cast: null pointer");
    if (in_type == out_type) {
        if (in == out) return BLADE_SUCCESS;   // the reference's bypass (cast/base.cc:45-49)
        // A plain copy, through the gather kernel (one workgroup per tile,
        // ~6.1 TB/s, against ~5.3 TB/s for hipMemcpyAsync device to device).
        return blade_gather(in, out, 1, n_elements * type_bytes(in_type), 1, 0, stream);
    }
    if (in == out) return fail("cast: in-place cast between different types");
    const uint64_t n = n_elements * type_components(in_type);
    const int os = scalar_of(out_type);
    switch (scalar_of(in_type)) {
        case BLADE_I8: return dispatch_cast_out<int8_t>(os, in, out, n, s);
        case BLADE_F16: return dispatch_cast_out<_Float16>(os, in, out, n, s);
        case BLADE_F32: return dispatch_cast_out<float>(os, in, out, n, s);
        case BLADE_F64: return dispatch_cast_out<double>(os, in, out, n, s);
        case BLADE_I32: return dispatch_cast_out<int32_t>(os, in, out, n, s);
    }
    return fail("cast: unsupported input type");
}

extern "C" int blade_polarizer(int type, const void* input, void* output, uint64_t n_elements,
                               blade_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (n_elements % 2) return fail("polarizer: number of elements must be even (P = 2)");
    if (n_elements == 0) return BLADE_SUCCESS;
    if (!input || !output) return fail("polarizer: null pointer");
    const uint64_t pairs = n_elements / 2;
    if (type == BLADE_CF32) {
        if ((uintptr_t)input % 16 || (uintptr_t)output % 16)
            return fail("polarizer: CF32 buffers must be 16-byte aligned");
        hipLaunchKernelGGL(polarizer_cf32_kernel, dim3(grid_for(pairs, 256)), dim3(256), 0, s,
                           static_cast<const float4*>(input), static_cast<float4*>(output), pairs);
        return check_launch("polarizer");
    }
    if (type == BLADE_CF16) {
        if ((uintptr_t)input % 16 || (uintptr_t)output % 16)
            return fail("polarizer: CF16 buffers must be 16-byte aligned");
        const uint64_t quads = pairs / 2;
        if (quads)
            hipLaunchKernelGGL(polarizer_cf16_kernel, dim3(grid_for(quads, 256)), dim3(256), 0, s,
                               static_cast<const halfx8*>(input), static_cast<halfx8*>(output),
                               quads);
        if (pairs % 2)
            hipLaunchKernelGGL(polarizer_cf16_tail_kernel, dim3(1), dim3(1), 0, s,
                               static_cast<const _Float16*>(input), static_cast<_Float16*>(output),
                               pairs - 1);
        return check_launch("polarizer");
    }
    return fail("polarizer: type must be CF32 or CF16");
}

extern "C" int blade_gather(const void* input, void* output, uint64_t height,
                            uint64_t width_bytes, uint64_t multiplier, uint64_t step,
                            blade_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (multiplier == 0) return fail("gather: multiplier must be more than zero");
    if (step >= multiplier) return fail("gather: step must be below the multiplier");
    if (height == 0 || width_bytes == 0) return BLADE_SUCCESS;
    if (!input || !output) return fail("gather: null pointer");
    const uintptr_t a = (uintptr_t)input | (uintptr_t)output | width_bytes;
    if (a % 16 == 0) return launch_gather<Bytes16>(input, output, height, width_bytes, multiplier, step, s);
    if (a % 8 == 0) return launch_gather<uint64_t>(input, output, height, width_bytes, multiplier, step, s);
    if (a % 4 == 0) return launch_gather<uint32_t>(input, output, height, width_bytes, multiplier, step, s);
    if (a % 2 == 0) return launch_gather<uint16_t>(input, output, height, width_bytes, multiplier, step, s);
    return launch_gather<uint8_t>(input, output, height, width_bytes, multiplier, step, s);
}

extern "C" int blade_permutation(const void* input, void* output, uint64_t elem_bytes,
                                 const uint64_t* in_shape4, const uint64_t* indexes4,
                                 blade_stream_t stream) {
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!in_shape4 || !indexes4) return fail("permutation: null shape");
    bool seen[4] = {false, false, false, false};
    for (int d = 0; d < 4; d++) {
        if (indexes4[d] > 3) return fail("permutation: index out of range");
        if (seen[indexes4[d]]) return fail("permutation: repeated index");
        seen[indexes4[d]] = true;
    }
    uint64_t in_stride[4];
    in_stride[3] = 1;
    for (int d = 2; d >= 0; d--) in_stride[d] = in_stride[d + 1] * in_shape4[d + 1];
    uint64_t shape[4], stride[4];
    uint64_t total = 1;
    for (int d = 0; d < 4; d++) {
        shape[d] = in_shape4[indexes4[d]];
        stride[d] = in_stride[indexes4[d]];
        total *= shape[d];
    }
    if (total == 0) return BLADE_SUCCESS;
    // Canonical form: output axes that are also contiguous and in order in
    // the input merge into one (AFTP -> ATPF is a transpose of [F][T*P]), and
    // unit axes drop; the result is padded back to four axes at the front.
    uint64_t ms[4], mst[4];
    int n = 0;
    for (int d = 0; d < 4; d++) {
        if (shape[d] == 1) continue;
        if (n > 0 && mst[n - 1] == stride[d] * shape[d]) {
            ms[n - 1] *= shape[d];
            mst[n - 1] = stride[d];
        } else {
            ms[n] = shape[d];
            mst[n] = stride[d];
            n++;
        }
    }
    if (!input || !output) return fail("permutation: null pointer");
    if (input == output) return fail("permutation: in-place permutation is not supported");
    if (n <= 1 && (elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8 ||
                   elem_bytes == 16))
        return blade_gather(input, output, 1, total * elem_bytes, 1, 0, stream);   // a plain copy
    // An innermost axis that stays innermost and spans at most 16 bytes joins
    // the element (ATFP moves 16-byte (X, Y) pairs): the rest is then a tiled
    // transpose of whole elements instead of a gather of short runs.  Every
    // other input stride is a multiple of that axis's size.
    uint64_t esize = elem_bytes;
    if (n >= 2 && mst[n - 1] == 1) {
        const uint64_t folded = esize * ms[n - 1];
        const bool pow2 = (esize & (esize - 1)) == 0 && (folded & (folded - 1)) == 0;
        if (pow2 && esize <= 16 && folded <= 16 && ((uintptr_t)input | (uintptr_t)output) % folded == 0) {
            for (int k = 0; k < n - 1; k++) mst[k] /= ms[n - 1];
            esize = folded;
            n--;
        }
    }
    Perm4 p;
    for (int d = 0; d < 4; d++) {
        const int k = d - (4 - n);
        p.out_shape[d] = k >= 0 ? ms[k] : 1;
        p.in_stride_of_out_dim[d] = k >= 0 ? mst[k] : 0;
    }
    switch (esize) {
        case 1: return launch_permutation<uint8_t>(input, output, p, s);
        case 2: return launch_permutation<uint16_t>(input, output, p, s);
        case 4: return launch_permutation<uint32_t>(input, output, p, s);
        case 8: return launch_permutation<uint64_t>(input, output, p, s);
        case 16: return launch_permutation<Bytes16>(input, output, p, s);
    }
    return fail("permutation: element size must be 1, 2, 4, 8 or 16 bytes");
}
