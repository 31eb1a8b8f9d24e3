// Frequency-domain Green's function summation (ffdas_greens):
//   out[b, t, f] = sum_s x[b, s, f] e^{i k_f r_ts} / r_ts,   r_ts = |dst_t - src_s|.
// Fewer than 5 events (greens_kernel): one lane per target, its position and per-source
// distance in registers; the source position and the x values of a (source, frequency
// chunk, event chunk) are the same for the whole wave, so they come through the scalar unit;
// each thread keeps FPT frequencies x BT events of complex sums on the VALU.  5 or more
// events (greens_mfma_kernel): the sums on the MFMA, see there.  FP32 throughout (the input
// may be stored as FP16).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <climits>

#include "internal.h"

namespace ffdas {
namespace {

constexpr int kGreensFPT = 8;     // frequencies per thread

template <typename XT>
__device__ __forceinline__ float2 load_x(const XT *p);
template <>
__device__ __forceinline__ float2 load_x<float2>(const float2 *p) { return *p; }
template <>
__device__ __forceinline__ float2 load_x<__half2>(const __half2 *p) { return __half22float2(*p); }

template <typename XT, int BT>
__global__ __launch_bounds__(256) void greens_kernel(const float *__restrict__ src,
                                                     const float *__restrict__ wav,
                                                     const XT *__restrict__ x,
                                                     const float *__restrict__ dst,
                                                     float2 *__restrict__ out, int nsrc, int nfreq,
                                                     int ndst, int B) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int f0 = blockIdx.y * kGreensFPT;
    const int b0 = blockIdx.z * BT;
    const int nf = min(kGreensFPT, nfreq - f0), nb = min(BT, B - b0);
    const bool valid = t < ndst;
    float px = 0, py = 0, pz = 0;
    if (valid) {
        px = dst[3 * (int64_t)t];
        py = dst[3 * (int64_t)t + 1];
        pz = dst[3 * (int64_t)t + 2];
    }
    float k[kGreensFPT];
#pragma unroll
    for (int f = 0; f < kGreensFPT; f++) k[f] = f < nf ? wav[f0 + f] : 0.0f;
    float ar[kGreensFPT][BT], ai[kGreensFPT][BT];
#pragma unroll
    for (int f = 0; f < kGreensFPT; f++)
#pragma unroll
        for (int b = 0; b < BT; b++) ar[f][b] = ai[f][b] = 0.0f;

    for (int s = 0; s < nsrc; s++) {
        // r = rsq * rsqrt(rsq), as the reference (greens_kernels.cuh:59-61)
        const float dx = px - src[3 * s], dy = py - src[3 * s + 1], dz = pz - src[3 * s + 2];
        const float rsq = dx * dx + dy * dy + dz * dz;
        const float rinv = rsqrtf(rsq);
        const float r = rsq * rinv;
        const XT *xs = x + ((int64_t)b0 * nsrc + s) * nfreq + f0;
#pragma unroll
        for (int f = 0; f < kGreensFPT; f++) {
            if (f < nf) {
                float sn, cs;
                __sincosf(k[f] * r, &sn, &cs);
                const float gr = cs * rinv, gi = sn * rinv;
#pragma unroll
                for (int b = 0; b < BT; b++) {
                    if (b < nb) {
                        const float2 v = load_x<XT>(xs + (int64_t)b * nsrc * nfreq + f);
                        ar[f][b] += gr * v.x - gi * v.y;
                        ai[f][b] += gr * v.y + gi * v.x;
                    }
                }
            }
        }
    }
    if (!valid) return;
#pragma unroll
    for (int b = 0; b < BT; b++) {
        if (b >= nb) break;
        float2 *o = out + ((int64_t)(b0 + b) * ndst + t) * nfreq + f0;
#pragma unroll
        for (int f = 0; f < kGreensFPT; f++)
            if (f < nf) o[f] = float2{ar[f][b], ai[f][b]};
    }
}

// Many events: the multiply-adds on the MFMA.  Per wave a 16-target x 16-event tile of
// kGreensFPW frequencies; per step of 4 sources, lane (jl, g) = (lane % 16, lane / 16)
// generates the Green's function of target t0 + jl and source s0 + g (the A operand of
// v_mfma_f32_16x16x4f32) and loads x of event b0 + jl, source s0 + g (the B operand); the
// complex product is four real MFMAs (A_r B_r - A_i B_i, A_r B_i + A_i B_r).  Output register
// v of lane (jl, g) holds target t0 + 4 g + v, event b0 + jl.  One sincos per (target,
// source, frequency) is all the VALU does per term.
constexpr int kGreensFPW = 8;

typedef float GV4 __attribute__((ext_vector_type(4)));

template <typename XT>
__global__ __launch_bounds__(256) void greens_mfma_kernel(const float *__restrict__ src,
                                                          const float *__restrict__ wav,
                                                          const XT *__restrict__ x,
                                                          const float *__restrict__ dst,
                                                          float2 *__restrict__ out, int nsrc,
                                                          int nfreq, int ndst, int B) {
    const int lane = threadIdx.x & 63, jl = lane & 15, g = lane >> 4;
    const int t0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    const int f0 = blockIdx.y * kGreensFPW;
    const int b0 = blockIdx.z * 16;
    const int nf = min(kGreensFPW, nfreq - f0);
    const int ta = t0 + jl;                  // this lane's target in the A operand
    const bool tok = ta < ndst;
    float px = 0, py = 0, pz = 0;
    if (tok) {
        px = dst[3 * (int64_t)ta];
        py = dst[3 * (int64_t)ta + 1];
        pz = dst[3 * (int64_t)ta + 2];
    }
    const int be = b0 + jl;                  // this lane's event in the B operand
    const bool bok = be < B;
    float k[kGreensFPW];
#pragma unroll
    for (int f = 0; f < kGreensFPW; f++) k[f] = f < nf ? wav[f0 + f] : 0.0f;
    GV4 cr[kGreensFPW], ci[kGreensFPW];
#pragma unroll
    for (int f = 0; f < kGreensFPW; f++) {
        cr[f] = GV4{0.0f, 0.0f, 0.0f, 0.0f};
        ci[f] = GV4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    for (int s0 = 0; s0 < nsrc; s0 += 4) {
        const int s = s0 + g;
        const bool sok = s < nsrc;
        const bool aok = tok && sok;
        float rinv = 0.0f, r = 0.0f;
        if (aok) {
            const float dx = px - src[3 * s], dy = py - src[3 * s + 1], dz = pz - src[3 * s + 2];
            const float rsq = dx * dx + dy * dy + dz * dz;
            rinv = rsqrtf(rsq);
            r = rsq * rinv;
        }
        const XT *xs = x + ((int64_t)(bok ? be : 0) * nsrc + (sok ? s : 0)) * nfreq + f0;
        const bool xok = bok && sok;
#pragma unroll
        for (int f = 0; f < kGreensFPW; f++) {
            float sn, cs;
            __sincosf(k[f] * r, &sn, &cs);
            const float ar = aok ? cs * rinv : 0.0f, ai = aok ? sn * rinv : 0.0f;
            const float2 v = (xok && f < nf) ? load_x<XT>(xs + f) : float2{0.0f, 0.0f};
            cr[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, v.x, cr[f], 0, 0, 0);
            cr[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai, v.y, cr[f], 0, 0, 0);
            ci[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, v.y, ci[f], 0, 0, 0);
            ci[f] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, v.x, ci[f], 0, 0, 0);
        }
    }
    if (!bok) return;
#pragma unroll
    for (int v = 0; v < 4; v++) {
        const int t = t0 + 4 * g + v;
        if (t >= ndst) continue;
        float2 *o = out + ((int64_t)be * ndst + t) * nfreq + f0;
#pragma unroll
        for (int f = 0; f < kGreensFPW; f++)
            if (f < nf) o[f] = float2{cr[f][v], ci[f][v]};
    }
}

template <typename XT>
ffdas_error_t launch(hipStream_t st, const float *src, const float *wav, const void *x,
                     const float *dst, void *out, int nsrc, int nfreq, int ndst, int B) {
    // each (target, source, frequency) sincos once where it pays (measured): the MFMA tile
    // from 5 events (one 16-event tile instead of two VALU passes), the 4-event VALU kernel
    // for 2-4, the 1-event one for 1
    if (B >= 5) {
        const dim3 g((unsigned)((ndst + 63) / 64), (unsigned)((nfreq + kGreensFPW - 1) / kGreensFPW),
                     (unsigned)((B + 15) / 16));
        ProfScope ps(st, "greens");
        hipLaunchKernelGGL((greens_mfma_kernel<XT>), g, dim3(256), 0, st, src, wav,
                           static_cast<const XT *>(x), dst, static_cast<float2 *>(out), nsrc,
                           nfreq, ndst, B);
        return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
    }
    const dim3 grid((unsigned)((ndst + 255) / 256), (unsigned)((nfreq + kGreensFPT - 1) / kGreensFPT),
                    1);
    ProfScope ps(st, "greens");
    if (B >= 2) {
        const dim3 g(grid.x, grid.y, (unsigned)((B + 3) / 4));
        hipLaunchKernelGGL((greens_kernel<XT, 4>), g, dim3(256), 0, st, src, wav,
                           static_cast<const XT *>(x), dst, static_cast<float2 *>(out), nsrc,
                           nfreq, ndst, B);
    } else {
        const dim3 g(grid.x, grid.y, (unsigned)B);
        hipLaunchKernelGGL((greens_kernel<XT, 1>), g, dim3(256), 0, st, src, wav,
                           static_cast<const XT *>(x), dst, static_cast<float2 *>(out), nsrc,
                           nfreq, ndst, B);
    }
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

bool packed(const ffdas_tensor_desc &d) {
    int64_t expect = 1;
    for (size_t k = d.dims.size(); k-- > 0;) {
        if (d.dims[k] > 1 && d.strides[k] != expect) return false;
        expect *= d.dims[k];
    }
    return true;
}

}  // namespace
}  // namespace ffdas

using namespace ffdas;

static ffdas_error_t greens_call(ffdas_handle_t handle, const float *srcpos,
                                      const float *wavenums, ffdas_tensor_desc_t x_desc,
                                      const void *x, const float *dstpos,
                                      ffdas_tensor_desc_t out_desc, void *out) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    if (!x_desc || !out_desc) return FFDAS_ERROR_NULL_POINTER;
    FFDAS_TRY(handle->check_device());
    const ffdas_tensor_desc &xd = *x_desc, &od = *out_desc;
    const bool half_in = xd.dtype == FFDAS_C_16F;
    if ((xd.dtype != FFDAS_C_32F && !half_in) || od.dtype != FFDAS_C_32F)
        return FFDAS_ERROR_UNSUPPORTED_TYPE;
    if (xd.dims.size() != 3) return FFDAS_ERROR_INVALID_DIMS;
    const int64_t B = xd.dims[0], nsrc = xd.dims[1], nfreq = xd.dims[2];
    if (od.dims.size() != 3 || od.dims[0] != B || od.dims[2] != nfreq)
        return FFDAS_ERROR_INVALID_DIMS;
    const int64_t ndst = od.dims[1];
    if (B > INT_MAX || nsrc > INT_MAX || nfreq > INT_MAX || ndst > INT_MAX)
        return FFDAS_ERROR_DIMS_TOO_LARGE;
    if (!packed(xd) || !packed(od)) return FFDAS_ERROR_NON_CONTIGUOUS;
    if (B * ndst * nfreq == 0) return FFDAS_SUCCESS;
    if (!srcpos || !wavenums || !x || !dstpos || !out) return FFDAS_ERROR_NULL_POINTER;
    if ((B + 3) / 4 > 65535 || (nfreq + kGreensFPT - 1) / kGreensFPT > 65535)
        return FFDAS_ERROR_DIMS_TOO_LARGE;
    return half_in ? launch<__half2>(handle->stream, srcpos, wavenums, x, dstpos, out, (int)nsrc,
                                     (int)nfreq, (int)ndst, (int)B)
                   : launch<float2>(handle->stream, srcpos, wavenums, x, dstpos, out, (int)nsrc,
                                    (int)nfreq, (int)ndst, (int)B);
}

extern "C" ffdas_error_t ffdas_greens(ffdas_handle_t handle, const float *srcpos,
                                      const float *wavenums, ffdas_tensor_desc_t x_desc,
                                      const void *x, const float *dstpos,
                                      ffdas_tensor_desc_t out_desc, void *out) {
    ffdas::CallScope cs(handle, "ffdas_greens");
    return cs.done(greens_call(handle, srcpos, wavenums, x_desc, x, dstpos, out_desc, out));
}
