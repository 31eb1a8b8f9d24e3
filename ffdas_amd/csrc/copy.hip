// ffdas_contiguous_copy: an n-D strided read with element-type conversion into a contiguous
// output, for every (input, output) type pair the reference dispatches
// (src/tensor_ops/contiguous_copy.cu:25-115; kernel src/tensor_ops/tensor_copy_kernels.cuh:15-47;
// host side src/tensor_ops/contiguous_copy_impl.cuh:16-57).
//
// Design for gfx950: the dims are first collapsed on the host (size-1 dims dropped, adjacent
// dims merged where the input is contiguous across them), leaving an outer index space and
// one inner run.  The shape travels by value in the kernel arguments (no device allocation or
// host-to-device copy per call, so the call can be captured in a graph).  Each workgroup
// iteration copies one 1024-element segment of one inner run: the outer coordinates are
// decoded once per segment (uniform), and every lane converts 4 elements 256 apart, so the
// writes are contiguous across the lanes, as are the reads whenever the input's inner stride
// is 1.  All indexing is 64-bit and strides may be negative (the input pointer addresses the
// element at index 0, as a DLPack view does).
//
// Conversions (cast<Tx, Ty> of the reference, src/core/type_utils.h:242-259): int16 to FP32
// exactly, FP32 to FP16 / BF16 and FP64 to FP32 rounded to nearest even, widening exactly;
// complex types convert each part.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "internal.h"

namespace ffdas {
namespace {

constexpr int kMaxCopyDims = 16;       // ffdas tensors have at most 16 dims
constexpr int kCopyThreads = 256;
constexpr int kCopyPerThread = 4;
constexpr int kCopySeg = kCopyThreads * kCopyPerThread;

struct CopyShape {
    int nd;                            // outer dims after collapsing
    int64_t dims[kMaxCopyDims];        // outer dims, outermost first
    int64_t xs[kMaxCopyDims];          // input strides of the outer dims [elements]
    int64_t inner;                     // length of the inner run
    int64_t xs_inner;                  // input stride along the inner run [elements]
    int64_t rows;                      // product of the outer dims
    int64_t segs;                      // segments per inner run
};

// Scalar component types.
template <class T> struct Comp;
template <> struct Comp<int16_t> {
    static __device__ __forceinline__ float to_f(int16_t v) { return (float)v; }
};
template <> struct Comp<_Float16> {
    static __device__ __forceinline__ float to_f(_Float16 v) { return (float)v; }
};
struct BF16 { uint16_t bits; };
template <> struct Comp<BF16> {
    static __device__ __forceinline__ float to_f(BF16 v) { return __uint_as_float((uint32_t)v.bits << 16); }
};
template <> struct Comp<float> {
    static __device__ __forceinline__ float to_f(float v) { return v; }
};

template <class S, class D> struct Conv {
    // through FP32: exact for every source type here but double
    static __device__ __forceinline__ D run(S v);
};
template <class S> struct Conv<S, float> {
    static __device__ __forceinline__ float run(S v) { return Comp<S>::to_f(v); }
};
template <class S> struct Conv<S, _Float16> {
    static __device__ __forceinline__ _Float16 run(S v) { return (_Float16)Comp<S>::to_f(v); }
};
template <class S> struct Conv<S, BF16> {
    static __device__ __forceinline__ BF16 run(S v) {
        uint32_t u = __float_as_uint(Comp<S>::to_f(v));
        BF16 r;
        if ((u & 0x7fffffffu) > 0x7f800000u) {           // NaN stays NaN
            r.bits = (uint16_t)((u >> 16) | 0x40u);
        } else {
            u += 0x7fffu + ((u >> 16) & 1u);                 // round to nearest even
            r.bits = (uint16_t)(u >> 16);
        }
        return r;
    }
};
template <> struct Conv<int16_t, int16_t> {
    static __device__ __forceinline__ int16_t run(int16_t v) { return v; }
};
template <> struct Conv<float, double> {
    static __device__ __forceinline__ double run(float v) { return (double)v; }
};
template <> struct Conv<double, double> {
    static __device__ __forceinline__ double run(double v) { return v; }
};
template <> struct Conv<double, float> {
    static __device__ __forceinline__ float run(double v) { return (float)v; }
};

// One element = CW components (CW = 2 for complex types).
template <class S, class D, int CW>
__global__ __launch_bounds__(kCopyThreads) void strided_copy_kernel(const S *__restrict__ x,
                                                                    D *__restrict__ y,
                                                                    CopyShape sh) {
    const int64_t total = sh.rows * sh.segs;
    for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
        const int64_t row = it / sh.segs;
        const int64_t seg = it - row * sh.segs;
        // input offset of the row start: decode the outer coordinates (uniform per block)
        int64_t rem = row, xo = 0;
        for (int d = sh.nd - 1; d >= 0; d--) {
            const int64_t c = rem % sh.dims[d];
            rem /= sh.dims[d];
            xo += c * sh.xs[d];
        }
        const int64_t i0 = seg * kCopySeg + threadIdx.x;
        const int64_t yo = row * sh.inner;
#pragma unroll
        for (int j = 0; j < kCopyPerThread; j++) {
            const int64_t i = i0 + j * kCopyThreads;     // consecutive lanes, consecutive elements
            if (i >= sh.inner) break;
            const S *src = x + (xo + i * sh.xs_inner) * CW;
            D *dst = y + (yo + i) * CW;
#pragma unroll
            for (int c = 0; c < CW; c++) dst[c] = Conv<S, D>::run(src[c]);
        }
    }
}

template <class S, class D, int CW>
ffdas_error_t launch_copy(hipStream_t st, const void *x, void *y, const CopyShape &sh) {
    const int64_t total = sh.rows * sh.segs;
    const unsigned blocks = (unsigned)std::min<int64_t>(total, 65536);
    ProfScope ps(st, "contiguous_copy");
    strided_copy_kernel<S, D, CW><<<blocks, kCopyThreads, 0, st>>>(static_cast<const S *>(x),
                                                                   static_cast<D *>(y), sh);
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

// Collapse the input's dims: drop size-1 dims, merge a dim into its inner neighbour where
// stride[outer] == stride[inner] * dims[inner].  The last remaining dim is the inner run.
bool collapse(const ffdas_tensor_desc &xd, CopyShape &sh) {
    std::vector<int64_t> d, s;
    for (size_t i = 0; i < xd.dims.size(); i++) {
        if (xd.dims[i] == 1) continue;
        if (!d.empty() && s.back() == xd.strides[i] * xd.dims[i]) {
            d.back() *= xd.dims[i];
            s.back() = xd.strides[i];
        } else {
            d.push_back(xd.dims[i]);
            s.push_back(xd.strides[i]);
        }
    }
    if (d.empty()) {          // a single element
        d.push_back(1);
        s.push_back(1);
    }
    if ((int)d.size() - 1 > kMaxCopyDims) return false;
    sh.nd = (int)d.size() - 1;
    sh.rows = 1;
    for (int i = 0; i < sh.nd; i++) {
        sh.dims[i] = d[i];
        sh.xs[i] = s[i];
        sh.rows *= d[i];
    }
    sh.inner = d.back();
    sh.xs_inner = s.back();
    sh.segs = (sh.inner + kCopySeg - 1) / kCopySeg;
    return true;
}

// (input type, output type) pairs of the reference's dispatch, plus the bfloat16 extension
// types, which mirror the FP16 rows.
ffdas_error_t dispatch(hipStream_t st, ffdas_datatype_t tx, ffdas_datatype_t ty, const void *x,
                       void *y, const CopyShape &sh, bool run) {
#define FFDAS_COPY_PAIR(TX, TY, S, D, CW) \
    if (tx == TX && ty == TY) return run ? launch_copy<S, D, CW>(st, x, y, sh) : FFDAS_SUCCESS;
    FFDAS_COPY_PAIR(FFDAS_R_16I, FFDAS_R_16I, int16_t, int16_t, 1)
    FFDAS_COPY_PAIR(FFDAS_R_16I, FFDAS_R_32F, int16_t, float, 1)
    FFDAS_COPY_PAIR(FFDAS_R_16F, FFDAS_R_16F, _Float16, _Float16, 1)
    FFDAS_COPY_PAIR(FFDAS_R_16F, FFDAS_R_32F, _Float16, float, 1)
    FFDAS_COPY_PAIR(FFDAS_R_32F, FFDAS_R_16F, float, _Float16, 1)
    FFDAS_COPY_PAIR(FFDAS_R_32F, FFDAS_R_32F, float, float, 1)
    FFDAS_COPY_PAIR(FFDAS_R_32F, FFDAS_R_64F, float, double, 1)
    FFDAS_COPY_PAIR(FFDAS_C_16I, FFDAS_C_16I, int16_t, int16_t, 2)
    FFDAS_COPY_PAIR(FFDAS_C_16I, FFDAS_C_32F, int16_t, float, 2)
    FFDAS_COPY_PAIR(FFDAS_C_16F, FFDAS_C_16F, _Float16, _Float16, 2)
    FFDAS_COPY_PAIR(FFDAS_C_16F, FFDAS_C_32F, _Float16, float, 2)
    FFDAS_COPY_PAIR(FFDAS_C_32F, FFDAS_C_16F, float, _Float16, 2)
    FFDAS_COPY_PAIR(FFDAS_C_32F, FFDAS_C_32F, float, float, 2)
    FFDAS_COPY_PAIR(FFDAS_C_32F, FFDAS_C_64F, float, double, 2)
    FFDAS_COPY_PAIR(FFDAS_R_64F, FFDAS_R_64F, double, double, 1)
    FFDAS_COPY_PAIR(FFDAS_R_64F, FFDAS_R_32F, double, float, 1)
    FFDAS_COPY_PAIR(FFDAS_C_64F, FFDAS_C_64F, double, double, 2)
    FFDAS_COPY_PAIR(FFDAS_C_64F, FFDAS_C_32F, double, float, 2)
    // extension: bfloat16 storage types
    FFDAS_COPY_PAIR(FFDAS_R_16BF, FFDAS_R_16BF, BF16, BF16, 1)
    FFDAS_COPY_PAIR(FFDAS_R_16BF, FFDAS_R_32F, BF16, float, 1)
    FFDAS_COPY_PAIR(FFDAS_R_32F, FFDAS_R_16BF, float, BF16, 1)
    FFDAS_COPY_PAIR(FFDAS_C_16BF, FFDAS_C_16BF, BF16, BF16, 2)
    FFDAS_COPY_PAIR(FFDAS_C_16BF, FFDAS_C_32F, BF16, float, 2)
    FFDAS_COPY_PAIR(FFDAS_C_32F, FFDAS_C_16BF, float, BF16, 2)
#undef FFDAS_COPY_PAIR
    return FFDAS_ERROR_UNSUPPORTED_TYPE;
}

}  // namespace
}  // namespace ffdas

using namespace ffdas;

// Validation in the reference's order (contiguous_copy.cu:14-22, 115; contiguous_copy_impl.cuh:
// 22-23): handle, device, descriptors, then the same dims and a contiguous output
// (FFDAS_ERROR_INVALID_DIMS), then the type pair (FFDAS_ERROR_UNSUPPORTED_TYPE).  Differences,
// all permissive: a size-1 output dim may have any stride; a tensor beyond 32-bit indexing is
// copied (64-bit indexing) instead of being rejected; a one-element tensor whose dims are all 1
// is copied (the reference kernel returns early only for zero dims, which descriptors cannot
// have).  Null data pointers are rejected with FFDAS_ERROR_NULL_POINTER when there is
// something to copy.
static ffdas_error_t contiguous_copy_call(ffdas_handle_t handle, ffdas_tensor_desc_t x_desc,
                                               const void *x, ffdas_tensor_desc_t out_desc,
                                               void *out) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    FFDAS_TRY(handle->check_device());
    if (!x_desc || !out_desc) return FFDAS_ERROR_NULL_POINTER;
    const ffdas_tensor_desc &xd = *x_desc, &od = *out_desc;
    if (xd.dims != od.dims || xd.strides.size() != xd.dims.size()) return FFDAS_ERROR_INVALID_DIMS;
    int64_t s = 1;
    for (size_t i = od.dims.size(); i-- > 0;) {
        if (od.dims[i] < 0) return FFDAS_ERROR_INVALID_DIMS;
        if (od.dims[i] > 1 && od.strides[i] != s) return FFDAS_ERROR_INVALID_DIMS;
        s *= od.dims[i];
    }
    CopyShape sh{};
    FFDAS_TRY(dispatch(handle->stream, xd.dtype, od.dtype, nullptr, nullptr, sh, false));
    if (xd.numel() == 0) return FFDAS_SUCCESS;
    if (!x || !out) return FFDAS_ERROR_NULL_POINTER;
    if (!collapse(xd, sh)) return FFDAS_ERROR_INVALID_DIMS;
    return dispatch(handle->stream, xd.dtype, od.dtype, x, out, sh, true);
}

extern "C" ffdas_error_t ffdas_contiguous_copy(ffdas_handle_t handle, ffdas_tensor_desc_t x_desc,
                                               const void *x, ffdas_tensor_desc_t out_desc,
                                               void *out) {
    ffdas::CallScope cs(handle, "ffdas_contiguous_copy");
    return cs.done(contiguous_copy_call(handle, x_desc, x, out_desc, out));
}
