// Device-side building blocks shared by the delay-and-sum kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "internal.h"

namespace ffdas {
namespace dev {

constexpr float kRadiusSqMin = 1e-7f;   // src/das/das_alg1_kernels.cuh:64 (float there)

// ---------------------------------------------------------------------------------------
// element access per storage class
// ---------------------------------------------------------------------------------------
template <Storage ST> struct Elem;
template <> struct Elem<Storage::F16> {
    using raw = _Float16;
    static __device__ __forceinline__ float real(raw v) { return (float)v; }
};
template <> struct Elem<Storage::BF16> {
    using raw = uint16_t;
    static __device__ __forceinline__ float real(raw v) { return __uint_as_float((uint32_t)v << 16); }
};
template <> struct Elem<Storage::F32> {
    using raw = float;
    static __device__ __forceinline__ float real(raw v) { return v; }
};
template <> struct Elem<Storage::F64> {
    using raw = double;
    static __device__ __forceinline__ double real(raw v) { return v; }
};

// Load the pair (x[k], x[k+1]) from a frame row.  `row` already points at element k of
// the row (complex: element = (re, im)).
template <Storage ST, class A>
__device__ __forceinline__ void load_pair(const typename Elem<ST>::raw *p, A &x0, A &x1) {
    x0 = (A)Elem<ST>::real(p[0]);
    x1 = (A)Elem<ST>::real(p[1]);
}
template <Storage ST, class A>
__device__ __forceinline__ void load_pair_c(const typename Elem<ST>::raw *p, A &x0r, A &x0i,
                                            A &x1r, A &x1i) {
    x0r = (A)Elem<ST>::real(p[0]);
    x0i = (A)Elem<ST>::real(p[1]);
    x1r = (A)Elem<ST>::real(p[2]);
    x1i = (A)Elem<ST>::real(p[3]);
}

// ---------------------------------------------------------------------------------------
// receive radius and delay
// ---------------------------------------------------------------------------------------
// A point split into an integer part and a fraction (both exact in float), so that
// differences of coordinates are exact integers plus exact small fractions.
struct SplitPt {
    float cx, cy, cz;  // rint(coordinate)
    float fx, fy, fz;  // coordinate - rint(coordinate)
    float x, y, z;     // coordinate
};

__device__ __forceinline__ SplitPt split_point(float x, float y, float z) {
#pragma clang fp contract(off)
    SplitPt s;
    s.x = x; s.y = y; s.z = z;
    s.cx = __builtin_rintf(x); s.cy = __builtin_rintf(y); s.cz = __builtin_rintf(z);
    s.fx = x - s.cx; s.fy = y - s.cy; s.fz = z - s.cz;
    return s;
}

struct Radius {
    float rh, rl;       // r = rh + rl
    float rinv;         // ~1/r
    float dx, dy, dz;   // voxel - channel
};

// Plain FP32: r = r2 * rsqrt(r2) as the reference does (das_alg1_kernels.cuh:64-66).  Every
// launcher instantiates the kernels with the compensated radius below (PRECISE = true); this
// form is the one the CPU oracle's FP32 build restates.
__device__ __forceinline__ Radius radius_fast(const SplitPt &y, const SplitPt &p) {
    Radius R;
    R.dx = y.x - p.x; R.dy = y.y - p.y; R.dz = y.z - p.z;
    float r2 = __builtin_fmaf(R.dx, R.dx, __builtin_fmaf(R.dy, R.dy, R.dz * R.dz));
    r2 = __builtin_fmaxf(r2, kRadiusSqMin);
    R.rinv = __builtin_amdgcn_rsqf(r2);
    R.rh = r2 * R.rinv;
    R.rl = 0.0f;
    return R;
}

// The error-free transformations below (TwoSum, exact products via fmaf) rely on every
// product and sum being rounded exactly where written, so floating-point contraction is
// switched off inside these functions; intended fused operations are explicit fmaf calls.
//
// Compensated FP32: the coordinate differences are formed exactly as integer + fraction,
// the squared radius as an exact 32-bit integer plus an FP32 remainder, and one Newton
// correction gives r = rh + rl with an error of a few 1e-8 samples over the whole range
// of delays (instead of ~1e-4 for a single FP32 value near 2000 samples).
__device__ __forceinline__ Radius radius_precise(const SplitPt &y, const SplitPt &p) {
#pragma clang fp contract(off)
    Radius R;
    const float dcx = y.cx - p.cx, dcy = y.cy - p.cy, dcz = y.cz - p.cz;   // exact integers
    const float dfx = y.fx - p.fx, dfy = y.fy - p.fy, dfz = y.fz - p.fz;   // |.| <= 1
    R.dx = dcx + dfx; R.dy = dcy + dfy; R.dz = dcz + dfz;
    const float lim = 32767.0f;
    const bool big = __builtin_fabsf(dcx) > lim || __builtin_fabsf(dcy) > lim ||
                     __builtin_fabsf(dcz) > lim;
    // integer part of |d|^2, exact in 32 bits for |dc| <= 32767
    const int ix = (int)__builtin_fminf(__builtin_fmaxf(dcx, -lim), lim);
    const int iy = (int)__builtin_fminf(__builtin_fmaxf(dcy, -lim), lim);
    const int iz = (int)__builtin_fminf(__builtin_fmaxf(dcz, -lim), lim);
    const uint32_t I = (uint32_t)(ix * ix) + (uint32_t)(iy * iy) + (uint32_t)(iz * iz);
    // remainder 2*dc*df + df^2 per axis
    const float F = __builtin_fmaf(dfx, __builtin_fmaf(2.0f, dcx, dfx),
                    __builtin_fmaf(dfy, __builtin_fmaf(2.0f, dcy, dfy),
                                   dfz * __builtin_fmaf(2.0f, dcz, dfz)));
    const float Ih = (float)(I & 0xFFFFFF00u);    // <= 24 significant bits: exact
    const float Il = (float)(I & 0xFFu);
    const float r2u = Ih + (Il + F);
    const float r2 = __builtin_fmaxf(r2u, kRadiusSqMin);
    R.rinv = __builtin_amdgcn_rsqf(r2);
    const float r0 = r2 * R.rinv;
    const float ph = r0 * r0;
    const float pl = __builtin_fmaf(r0, r0, -ph);
    float res = (Ih - ph);                         // exact (Sterbenz)
    res = (res - pl) + (Il + F);
    float rl = res * (0.5f * R.rinv);
    if (r2u < kRadiusSqMin || big) rl = 0.0f;
    if (big) {  // far outside any imaging geometry: plain FP32
        const float q = __builtin_fmaf(R.dx, R.dx, __builtin_fmaf(R.dy, R.dy, R.dz * R.dz));
        R.rinv = __builtin_amdgcn_rsqf(q);
        R.rh = q * R.rinv;
        R.rl = 0.0f;
        return R;
    }
    R.rh = r0;
    R.rl = rl;
    return R;
}

struct Tap {
    int k;      // floor(tau)
    float kf;   // k as float
    float t;    // tau - k in [0, 1)
};

// tau = scale * (ofs + r)
template <bool PRECISE>
__device__ __forceinline__ Tap make_tap(float ofs, const Radius &R, float scale) {
#pragma clang fp contract(off)
    Tap tp;
    if constexpr (!PRECISE) {
        float tau = ofs + R.rh;
        tau *= scale;                      // exact for scale == 1
        tp.kf = __builtin_floorf(tau);
        tp.t = tau - tp.kf;
    } else {
        // TwoSum(ofs, rh)
        float th = ofs + R.rh;
        float bp = th - ofs;
        float ap = th - bp;
        float tl = (ofs - ap) + (R.rh - bp);
        tl += R.rl;
        // times scale, the product error kept in tl; for scale == 1 this leaves th and tl
        // unchanged (sl = 0), so it needs no branch
        const float sh = scale * th;
        const float sl = __builtin_fmaf(scale, th, -sh);
        tl = __builtin_fmaf(scale, tl, sl);
        th = sh;
        float kf = __builtin_floorf(th);
        float t = (th - kf) + tl;
        float c = __builtin_floorf(t);
        tp.kf = kf + c;
        tp.t = t - c;
    }
    tp.k = (int)tp.kf;
    return tp;
}

// ---------------------------------------------------------------------------------------
// phase rotation e^{i wavenum tau}
// ---------------------------------------------------------------------------------------
// sin/cos of 2*pi*u for |u| <= 0.5 by quadrant reduction and Taylor polynomials on
// [-pi/4, pi/4] (truncation below 2e-9).
__device__ __forceinline__ void sincos_rev_poly(float u, float &s, float &c) {
#pragma clang fp contract(off)
    const float q = __builtin_rintf(4.0f * u);
    const float v = __builtin_fmaf(-0.25f, q, u);
    const float x = v * 6.28318530717958648f;
    const float x2 = x * x;
    float ps = __builtin_fmaf(x2, 2.7557319223985893e-6f, -1.9841269841269841e-4f);
    ps = __builtin_fmaf(x2, ps, 8.3333333333333333e-3f);
    ps = __builtin_fmaf(x2, ps, -1.6666666666666667e-1f);
    const float sn = __builtin_fmaf(x * x2, ps, x);
    float pc = __builtin_fmaf(x2, 2.4801587301587302e-5f, -1.3888888888888889e-3f);
    pc = __builtin_fmaf(x2, pc, 4.1666666666666667e-2f);
    pc = __builtin_fmaf(x2, pc, -0.5f);
    const float cs = __builtin_fmaf(x2, pc, 1.0f);
    const int qi = ((int)q) & 3;
    s = (qi == 0) ? sn : (qi == 1) ? cs : (qi == 2) ? -sn : -cs;
    c = (qi == 0) ? cs : (qi == 1) ? -sn : (qi == 2) ? -cs : sn;
}

// Phase of wavenum*tau in revolutions, reduced to [-0.5, 0.5].
// rev_h + rev_l = wavenum / (2 pi) (double-float), tau = kf + t.
template <bool PRECISE>
__device__ __forceinline__ void phase(float rev_h, float rev_l, const Tap &tp, float &s, float &c) {
#pragma clang fp contract(off)
    if constexpr (PRECISE) {
        const float p = rev_h * tp.kf;
        const float pe = __builtin_fmaf(rev_h, tp.kf, -p);
        float u = p - __builtin_rintf(p);
        u += pe + __builtin_fmaf(rev_l, tp.kf, (rev_h + rev_l) * tp.t);
        u -= __builtin_rintf(u);
        sincos_rev_poly(u, s, c);
    } else {
        float u = rev_h * (tp.kf + tp.t);
        u -= __builtin_rintf(u);
        s = __builtin_amdgcn_sinf(u);
        c = __builtin_amdgcn_cosf(u);
    }
}

// ---------------------------------------------------------------------------------------
// voxel tiling: a wave covers a compact (ta x th x tw) tile of the (A, H, W) voxel grid
// ---------------------------------------------------------------------------------------
struct Tiling {
    int A, H, W;        // spatial view (outer dims folded into A)
    int ta, th, tw;     // wave tile, ta*th*tw == 64
    int nta, nth, ntw;  // tile counts
    int xcd;            // 1: blocks regrouped so that each XCD takes a contiguous range
};

// Workgroups are dispatched to the eight XCDs in turn (linear block id mod 8).  With xcd
// set, linear id b is read as block (b mod 8) * (nb / 8) + b / 8 of the grid's order, so the
// blocks of one XCD form a contiguous range of tiles and share that XCD's L2.  A bijection
// on the complete rounds of 8; the last nb mod 8 blocks keep their ids.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int full = nb & ~7;
    if (b >= full) return b;
    return (b & 7) * (full >> 3) + (b >> 3);
}

__device__ __forceinline__ int voxel_of(const Tiling &tl, int tile, int lane, bool &valid) {
    const int ntiles = tl.nta * tl.nth * tl.ntw;
    valid = tile < ntiles;
    const int iw = tile % tl.ntw;
    const int rest = tile / tl.ntw;
    const int ih = rest % tl.nth;
    const int ia = rest / tl.nth;
    const int lw = lane % tl.tw;
    const int lh = (lane / tl.tw) % tl.th;
    const int la = lane / (tl.tw * tl.th);
    const int w = iw * tl.tw + lw, h = ih * tl.th + lh, a = ia * tl.ta + la;
    valid = valid && w < tl.W && h < tl.H && a < tl.A;
    return (a * tl.H + h) * tl.W + w;
}

// Axial voxel order for the channel-row scatter: the spatial grid is viewed as (D0, R) with
// D0 its first non-unit dimension (depth, in the (z, y, x) grids the reference's examples
// build) and R the product of the others.  Voxel n = d0 * R + r is visited as
// q = r * D0 + d0, and a wave takes 2^lg consecutive d0 for 64 >> lg values of r: along depth
// the delay grows by several samples per voxel, so the lanes of one scatter instruction hit
// distinct LDS samples (no same-address serialisation).
struct Axial {
    int D0, R;
    int lg;   // log2 of the lanes along d0
};

// Problem geometry passed to every kernel by value.
struct Geo {
    const float *srcpos;   // (C,3)
    const float *srcdir;   // (C,4) or null
    const float *dstpos;   // (N,3)
    const float *offsets;  // (K,N)
    const float *weights;  // (K,N)
    const int *sparse;     // (K,N) or null
    int C, T, S, K, B, N;
    float scale;
    float rev_h, rev_l;    // wavenum / 2pi
    float wavenum;
};

}  // namespace dev
}  // namespace ffdas
