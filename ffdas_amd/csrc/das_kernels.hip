// gfx950 delay-and-sum kernels: forward gather, adjoint (parameter gather + channel-row
// scatter), and the cast/relayout copy.
//
// Forward (semantics: src/das/das_alg1_kernels.cuh:41-106).  One lane per voxel; a wave
// covers a compact 64-voxel tile so the 64 gathers of one (channel, transmit) row touch
// a narrow sample window (few cache lines per load instruction).  The receive radius
// r(voxel, channel) is computed once per channel and reused for a chunk of KC transmit
// slots held in registers; a tile of FT frames is accumulated per lane so the delay
// arithmetic is amortised over FT interpolations.  Channel positions are uniform across
// the wave (scalar loads).
//
// Adjoint w.r.t. offsets / weights / scale: same gather structure, reducing
// grad_out * (x[k+1] - x[k]) and grad_out * lerp per (slot, voxel).
//
// Adjoint w.r.t. x: a workgroup owns one channel-data row (channel m, transmit s) for a
// chunk of FS frames, scatters every voxel's two interpolation weights into that row in
// LDS with LDS float atomics, then writes the row once (no global atomics).
#include "das_device.h"

namespace ffdas {
namespace dev {

template <bool ACC64> using AccT = typename std::conditional<ACC64, double, float>::type;

// ---------------------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------------------
template <Storage ST, bool ACC64, bool CPLX, bool PRECISE, int FT, int KC>
__global__ __launch_bounds__(256) void das_fwd_kernel(Geo g, Tiling tl, const void *__restrict__ xv,
                                                      int64_t xs_b, int64_t xs_c, int64_t xs_t,
                                                      void *__restrict__ outv, int64_t os_b,
                                                      double beta_re, double beta_im) {
    using A = AccT<ACC64>;
    using raw = typename Elem<ST>::raw;
    constexpr int CW = CPLX ? 2 : 1;
    const raw *__restrict__ x = static_cast<const raw *>(xv);

    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    bool valid;
    const int n = voxel_of(tl, tile, lane, valid);
    if (!valid) return;

    const int b0 = blockIdx.y * FT;
    const SplitPt yp = split_point(g.dstpos[3 * n], g.dstpos[3 * n + 1], g.dstpos[3 * n + 2]);

    A acc[FT][CW];
#pragma unroll
    for (int f = 0; f < FT; f++)
#pragma unroll
        for (int c = 0; c < CW; c++) acc[f][c] = A(0);

    // frame base pointers (uniform); frames past B re-read the last frame (never stored)
    const raw *fb[FT];
#pragma unroll
    for (int f = 0; f < FT; f++) {
        const int b = min(b0 + f, g.B - 1);
        fb[f] = x + (int64_t)b * xs_b * CW;
    }
    const bool unit = g.unit_scale != 0;

    for (int kc = 0; kc < g.K; kc += KC) {
        float ofs[KC], wt[KC];
        int sq[KC];
#pragma unroll
        for (int j = 0; j < KC; j++) {
            const int jj = kc + j;
            if (jj < g.K) {
                const int64_t o = (int64_t)jj * g.N + n;
                ofs[j] = g.offsets[o];
                wt[j] = g.weights[o];
                sq[j] = g.sparse ? g.sparse[o] : jj;
            } else {
                ofs[j] = 0.0f;
                wt[j] = 0.0f;
                sq[j] = -1;
            }
        }
        for (int m = 0; m < g.C; m++) {
            const SplitPt pp = split_point(g.srcpos[3 * m], g.srcpos[3 * m + 1], g.srcpos[3 * m + 2]);
            const Radius R = PRECISE ? radius_precise(yp, pp) : radius_fast(yp, pp);
            bool dok = true;
            if (g.srcdir) {
                const float *sd = g.srcdir + 4 * m;
                dok = (R.dx * sd[0] + R.dy * sd[1] + R.dz * sd[2]) * R.rinv > sd[3];
            }
            const uint32_t rowc = (uint32_t)((int64_t)m * xs_c);
#pragma unroll
            for (int j = 0; j < KC; j++) {
                if (kc + j >= g.K) break;
                const Tap tp = make_tap<PRECISE>(ofs[j], R, g.scale, unit);
                const bool in = dok && (uint32_t)tp.k < (uint32_t)(g.S - 1) &&
                                (uint32_t)sq[j] < (uint32_t)g.T;
                if (in) {
                    const uint32_t off = rowc + (uint32_t)((int64_t)sq[j] * xs_t) + (uint32_t)tp.k;
                    if constexpr (!CPLX) {
                        const A c = (A)wt[j] * (A)tp.t;
                        const A a = (A)wt[j] - c;
#pragma unroll
                        for (int f = 0; f < FT; f++) {
                            A x0, x1;
                            load_pair<ST>(fb[f] + off, x0, x1);
                            acc[f][0] = __builtin_fma(c, x1, __builtin_fma(a, x0, acc[f][0]));
                        }
                    } else {
                        float sn, cs;
                        phase<PRECISE>(g.rev_h, g.rev_l, tp, sn, cs);
                        const float wr = wt[j] * cs, wi = wt[j] * sn;
                        const A t = (A)tp.t;
#pragma unroll
                        for (int f = 0; f < FT; f++) {
                            A x0r, x0i, x1r, x1i;
                            load_pair_c<ST>(fb[f] + 2 * (size_t)off, x0r, x0i, x1r, x1i);
                            const A lr = x0r + t * (x1r - x0r);
                            const A li = x0i + t * (x1i - x0i);
                            acc[f][0] += (A)wr * lr - (A)wi * li;
                            acc[f][1] += (A)wr * li + (A)wi * lr;
                        }
                    }
                }
            }
        }
    }

    A *out = static_cast<A *>(outv);
    const bool zero_beta = beta_re == 0.0 && beta_im == 0.0;
#pragma unroll
    for (int f = 0; f < FT; f++) {
        const int b = b0 + f;
        if (b >= g.B) break;
        A *o = out + ((int64_t)b * os_b + n) * CW;
        if constexpr (!CPLX) {
            o[0] = zero_beta ? acc[f][0] : (A)beta_re * o[0] + acc[f][0];
        } else {
            if (zero_beta) {
                o[0] = acc[f][0];
                o[1] = acc[f][1];
            } else {
                const A orr = o[0], oi = o[1];
                o[0] = (A)beta_re * orr - (A)beta_im * oi + acc[f][0];
                o[1] = (A)beta_re * oi + (A)beta_im * orr + acc[f][1];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// adjoint: offsets / weights / scale
// ---------------------------------------------------------------------------------------
template <bool ACC64>
__device__ __forceinline__ AccT<ACC64> wave_sum(AccT<ACC64> v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <Storage ST, bool ACC64, bool CPLX, bool PRECISE, int FT, int KC>
__global__ __launch_bounds__(256) void das_bwd_param_kernel(
    Geo g, Tiling tl, const void *__restrict__ xv, int64_t xs_b, int64_t xs_c, int64_t xs_t,
    const void *__restrict__ goutv, int64_t gs_b, float *__restrict__ gofs,
    float *__restrict__ gwt, float *__restrict__ gscale, int accumulate) {
    using A = AccT<ACC64>;
    using raw = typename Elem<ST>::raw;
    constexpr int CW = CPLX ? 2 : 1;
    const raw *__restrict__ x = static_cast<const raw *>(xv);
    const A *__restrict__ gout = static_cast<const A *>(goutv);

    __shared__ A red[8];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tile = blockIdx.x * (blockDim.x >> 6) + wave;
    bool valid;
    const int n = voxel_of(tl, tile, lane, valid);
    A scale_part = A(0);

    if (valid) {
        const int b0 = blockIdx.y * FT;
        const SplitPt yp = split_point(g.dstpos[3 * n], g.dstpos[3 * n + 1], g.dstpos[3 * n + 2]);
        A gv[FT][CW];
        const raw *fb[FT];
#pragma unroll
        for (int f = 0; f < FT; f++) {
            const int b = b0 + f;
            const int bc = min(b, g.B - 1);
            fb[f] = x + (int64_t)bc * xs_b * CW;
#pragma unroll
            for (int c = 0; c < CW; c++)
                gv[f][c] = (b < g.B) ? gout[((int64_t)b * gs_b + n) * CW + c] : A(0);
        }
        const bool unit = g.unit_scale != 0;

        for (int kc = 0; kc < g.K; kc += KC) {
            float ofs[KC], wt[KC];
            int sq[KC];
            A d_tau[KC], d_w[KC];
#pragma unroll
            for (int j = 0; j < KC; j++) {
                const int jj = kc + j;
                d_tau[j] = A(0);
                d_w[j] = A(0);
                if (jj < g.K) {
                    const int64_t o = (int64_t)jj * g.N + n;
                    ofs[j] = g.offsets[o];
                    wt[j] = g.weights[o];
                    sq[j] = g.sparse ? g.sparse[o] : jj;
                } else {
                    ofs[j] = 0.0f;
                    wt[j] = 0.0f;
                    sq[j] = -1;
                }
            }
            for (int m = 0; m < g.C; m++) {
                const SplitPt pp =
                    split_point(g.srcpos[3 * m], g.srcpos[3 * m + 1], g.srcpos[3 * m + 2]);
                const Radius R = PRECISE ? radius_precise(yp, pp) : radius_fast(yp, pp);
                bool dok = true;
                if (g.srcdir) {
                    const float *sd = g.srcdir + 4 * m;
                    dok = (R.dx * sd[0] + R.dy * sd[1] + R.dz * sd[2]) * R.rinv > sd[3];
                }
                const uint32_t rowc = (uint32_t)((int64_t)m * xs_c);
#pragma unroll
                for (int j = 0; j < KC; j++) {
                    if (kc + j >= g.K) break;
                    const Tap tp = make_tap<PRECISE>(ofs[j], R, g.scale, unit);
                    const bool in = dok && (uint32_t)tp.k < (uint32_t)(g.S - 1) &&
                                    (uint32_t)sq[j] < (uint32_t)g.T;
                    if (in) {
                        const uint32_t off =
                            rowc + (uint32_t)((int64_t)sq[j] * xs_t) + (uint32_t)tp.k;
                        const A t = (A)tp.t;
                        if constexpr (!CPLX) {
                            A sd_ = A(0), sl_ = A(0);
#pragma unroll
                            for (int f = 0; f < FT; f++) {
                                A x0, x1;
                                load_pair<ST>(fb[f] + off, x0, x1);
                                const A d = x1 - x0;
                                sd_ += gv[f][0] * d;
                                sl_ += gv[f][0] * (x0 + t * d);
                            }
                            const A dt = (A)wt[j] * sd_;
                            d_tau[j] += dt;
                            d_w[j] += sl_;
                            scale_part += dt * ((A)ofs[j] + (A)R.rh + (A)R.rl);
                        } else {
                            // Q = sum conj(g) (x1 - x0), L = sum conj(g) lerp
                            A qr = 0, qi = 0, lr = 0, li = 0;
#pragma unroll
                            for (int f = 0; f < FT; f++) {
                                A x0r, x0i, x1r, x1i;
                                load_pair_c<ST>(fb[f] + 2 * (size_t)off, x0r, x0i, x1r, x1i);
                                const A dr = x1r - x0r, di = x1i - x0i;
                                const A pr = x0r + t * dr, pi = x0i + t * di;
                                const A gr = gv[f][0], gi = gv[f][1];
                                qr += gr * dr + gi * di;
                                qi += gr * di - gi * dr;
                                lr += gr * pr + gi * pi;
                                li += gr * pi - gi * pr;
                            }
                            float sn, cs;
                            phase<PRECISE>(g.rev_h, g.rev_l, tp, sn, cs);
                            const A er = (A)cs, ei = (A)sn, kap = (A)g.wavenum;
                            // Re(e (Q + i kap L)), Re(e L)
                            const A zr = qr - kap * li, zi = qi + kap * lr;
                            const A dt = (A)wt[j] * (er * zr - ei * zi);
                            d_tau[j] += dt;
                            d_w[j] += er * lr - ei * li;
                            scale_part += dt * ((A)ofs[j] + (A)R.rh + (A)R.rl);
                        }
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < KC; j++) {
                const int jj = kc + j;
                if (jj >= g.K) break;
                const int64_t o = (int64_t)jj * g.N + n;
                const float vo = (float)(d_tau[j] * (A)g.scale);
                const float vw = (float)d_w[j];
                if (accumulate) {
                    if (gofs) atomicAdd(gofs + o, vo);
                    if (gwt) atomicAdd(gwt + o, vw);
                } else {
                    if (gofs) gofs[o] = vo;
                    if (gwt) gwt[o] = vw;
                }
            }
        }
    }
    if (gscale) {
        A v = wave_sum<ACC64>(scale_part);
        if (lane == 0) red[wave] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            A s = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
            atomicAdd(gscale, (float)s);
        }
    }
}

// ---------------------------------------------------------------------------------------
// adjoint: x (channel-row scatter through LDS)
// ---------------------------------------------------------------------------------------
// grid.x = C*T rows (m = row / T, s = row % T), grid.y = frame chunks of FS.
// LDS holds the row as [k][f] with a padded frame stride fsp = FS|1 (odd: conflict-free
// row write-back).  Voxels are visited in tile order (tiling of the forward) so that lanes
// of one wave fall on different samples.
template <bool ACC64, bool CPLX, bool PRECISE>
__global__ __launch_bounds__(512) void das_bwd_x_kernel(Geo g, Tiling tl, const void *__restrict__ goutv,
                                                        int64_t gs_b, void *__restrict__ gxv,
                                                        int64_t gx_b, int64_t gx_c, int64_t gx_t,
                                                        int FS) {
    using A = AccT<ACC64>;
    constexpr int CW = CPLX ? 2 : 1;
    const int FSP = FS | 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    A *lds = reinterpret_cast<A *>(lds_raw);
    const A *__restrict__ gout = static_cast<const A *>(goutv);
    A *__restrict__ gx = static_cast<A *>(gxv);

    const int row = blockIdx.x;
    const int m = row / g.T;
    const int s = row - m * g.T;
    const int b0 = blockIdx.y * FS;
    const int nf = min(FS, g.B - b0);
    const int nlds = g.S * FSP * CW;
    for (int i = threadIdx.x; i < nlds; i += blockDim.x) lds[i] = A(0);
    __syncthreads();

    const SplitPt pp = split_point(g.srcpos[3 * m], g.srcpos[3 * m + 1], g.srcpos[3 * m + 2]);
    float sdx = 0, sdy = 0, sdz = 0, sdw = 0;
    if (g.srcdir) {
        sdx = g.srcdir[4 * m]; sdy = g.srcdir[4 * m + 1]; sdz = g.srcdir[4 * m + 2];
        sdw = g.srcdir[4 * m + 3];
    }
    const bool unit = g.unit_scale != 0;
    const int lane = threadIdx.x & 63;
    const int waves = blockDim.x >> 6;
    const int ntiles = tl.nta * tl.nth * tl.ntw;
    for (int tile = threadIdx.x >> 6; tile < ntiles; tile += waves) {
        bool valid;
        const int n = voxel_of(tl, tile, lane, valid);
        if (!valid) continue;
        const SplitPt yp = split_point(g.dstpos[3 * n], g.dstpos[3 * n + 1], g.dstpos[3 * n + 2]);
        const Radius R = PRECISE ? radius_precise(yp, pp) : radius_fast(yp, pp);
        if (g.srcdir && !((R.dx * sdx + R.dy * sdy + R.dz * sdz) * R.rinv > sdw)) continue;
        const int jlo = g.sparse ? 0 : s;
        const int jhi = g.sparse ? g.K : s + 1;
        for (int j = jlo; j < jhi; j++) {
            const int64_t o = (int64_t)j * g.N + n;
            if (g.sparse && g.sparse[o] != s) continue;
            const Tap tp = make_tap<PRECISE>(g.offsets[o], R, g.scale, unit);
            if ((uint32_t)tp.k >= (uint32_t)(g.S - 1)) continue;
            const float w = g.weights[o];
            A *l0 = lds + (size_t)tp.k * FSP * CW;
            A *l1 = l0 + FSP * CW;
            const A t = (A)tp.t, omt = A(1) - (A)tp.t;
            if constexpr (!CPLX) {
                const A w0 = (A)w * omt, w1 = (A)w * t;
                for (int f = 0; f < nf; f++) {
                    const A gvv = gout[(int64_t)(b0 + f) * gs_b + n];
                    atomicAdd(l0 + f, w0 * gvv);
                    atomicAdd(l1 + f, w1 * gvv);
                }
            } else {
                float sn, cs;
                phase<PRECISE>(g.rev_h, g.rev_l, tp, sn, cs);
                const A cr = (A)(w * cs), ci = (A)(w * sn);
                for (int f = 0; f < nf; f++) {
                    const A *gp = gout + ((int64_t)(b0 + f) * gs_b + n) * 2;
                    const A gr = gp[0], gi = gp[1];
                    // conj(coef) * g
                    const A vr = cr * gr + ci * gi, vi = cr * gi - ci * gr;
                    atomicAdd(l0 + 2 * f, vr * omt);
                    atomicAdd(l0 + 2 * f + 1, vi * omt);
                    atomicAdd(l1 + 2 * f, vr * t);
                    atomicAdd(l1 + 2 * f + 1, vi * t);
                }
            }
        }
    }
    __syncthreads();
    // write back: consecutive threads -> consecutive samples
    for (int f = 0; f < nf; f++) {
        A *dst = gx + ((int64_t)(b0 + f) * gx_b + (int64_t)m * gx_c + (int64_t)s * gx_t) * CW;
        for (int k = threadIdx.x; k < g.S; k += blockDim.x) {
#pragma unroll
            for (int c = 0; c < CW; c++) dst[(size_t)k * CW + c] = lds[((size_t)k * FSP + f) * CW + c];
        }
    }
}

// ---------------------------------------------------------------------------------------
// fill (graph-safe replacement for memset of small outputs)
// ---------------------------------------------------------------------------------------
__global__ void fill_zero_kernel(float *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n; k += stride) p[k] = 0.0f;
}

// ---------------------------------------------------------------------------------------
// cast / relayout: strided ([B,]C,T,S) of type SRC -> contiguous (B,C,T,S) of type DST
// ---------------------------------------------------------------------------------------
template <class D> __device__ __forceinline__ D convert_to(float v);
template <> __device__ __forceinline__ float convert_to<float>(float v) { return v; }
template <> __device__ __forceinline__ _Float16 convert_to<_Float16>(float v) { return (_Float16)v; }
template <> __device__ __forceinline__ uint16_t convert_to<uint16_t>(float v) {
    // round to nearest even; NaN stays NaN
    uint32_t u = __float_as_uint(v);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <class SRC, class DST>
__global__ void cast_rows_kernel(const SRC *__restrict__ src, int64_t sb, int64_t sc, int64_t st,
                                 DST *__restrict__ dst, int C, int T, int S, int cw) {
    // one block row = one (b, c, t) row of S*cw scalars
    const int64_t r = blockIdx.x;
    const int t = (int)(r % T);
    const int c = (int)((r / T) % C);
    const int64_t b = r / ((int64_t)T * C);
    const SRC *sp = src + (b * sb + (int64_t)c * sc + (int64_t)t * st) * cw;
    DST *dp = dst + r * (int64_t)S * cw;
    const int n = S * cw;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        if constexpr (std::is_same<SRC, DST>::value)
            dp[i] = sp[i];
        else if constexpr (std::is_same<DST, double>::value)
            dp[i] = (double)sp[i];
        else
            dp[i] = convert_to<DST>((float)sp[i]);
    }
}

}  // namespace dev
}  // namespace ffdas

// =======================================================================================
// host launchers
// =======================================================================================
namespace ffdas {

namespace {

constexpr int kKC = 8;
constexpr int kLdsLimit = 160 * 1024;   // gfx950 LDS per workgroup

int frame_tile_for(int B) {
    if (B >= 16) return 16;
    if (B >= 8) return 8;
    if (B >= 4) return 4;
    if (B >= 2) return 2;
    return 1;
}

dev::Tiling make_tiling(const std::vector<int64_t> &spatial) {
    std::vector<int64_t> d;
    for (int64_t v : spatial)
        if (v > 1) d.push_back(v);
    dev::Tiling t{};
    t.A = 1; t.H = 1; t.W = 1;
    if (d.size() >= 1) t.W = (int)d.back();
    if (d.size() >= 2) t.H = (int)d[d.size() - 2];
    for (size_t i = 0; i + 2 < d.size(); i++) t.A *= (int)d[i];
    if (d.size() >= 3) { t.ta = 4; t.th = 4; t.tw = 4; }
    else if (d.size() == 2) { t.ta = 1; t.th = 8; t.tw = 8; }
    else { t.ta = 1; t.th = 1; t.tw = 64; }
    t.nta = (t.A + t.ta - 1) / t.ta;
    t.nth = (t.H + t.th - 1) / t.th;
    t.ntw = (t.W + t.tw - 1) / t.tw;
    return t;
}

dev::Geo make_geo(const DasGeomArgs &a, const DasProblem &p) {
    dev::Geo g{};
    g.srcpos = a.srcpos; g.srcdir = a.srcdir; g.dstpos = a.dstpos;
    g.offsets = a.offsets; g.weights = a.weights; g.sparse = p.sparse ? a.sparse_indices : nullptr;
    g.C = p.C; g.T = p.T; g.S = p.S; g.K = p.K; g.B = p.B; g.N = (int)p.N;
    g.scale = a.scale;
    g.unit_scale = a.scale == 1.0f;
    g.wavenum = a.wavenum;
    const double rev = (double)a.wavenum / (2.0 * 3.14159265358979323846);
    g.rev_h = (float)rev;
    g.rev_l = (float)(rev - (double)g.rev_h);
    return g;
}

template <Storage ST, bool ACC64, bool CPLX, int FT>
hipError_t launch_fwd_ft(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl, const XView &xv,
                         void *out, int64_t os_b, double br, double bi) {
    constexpr bool PRECISE = (ST == Storage::F32 || ST == Storage::F64);
    const int ntiles = tl.nta * tl.nth * tl.ntw;
    dim3 grid((ntiles + 3) / 4, (g.B + FT - 1) / FT);
    dev::das_fwd_kernel<ST, ACC64, CPLX, PRECISE, FT, kKC><<<grid, 256, 0, st>>>(
        g, tl, xv.ptr, xv.sb, xv.sc, xv.st, out, os_b, br, bi);
    return hipGetLastError();
}

template <Storage ST, bool ACC64, bool CPLX>
hipError_t launch_fwd_t(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl, const XView &xv,
                        void *out, int64_t os_b, double br, double bi) {
    switch (frame_tile_for(g.B)) {
        case 16: return launch_fwd_ft<ST, ACC64, CPLX, 16>(st, g, tl, xv, out, os_b, br, bi);
        case 8: return launch_fwd_ft<ST, ACC64, CPLX, 8>(st, g, tl, xv, out, os_b, br, bi);
        case 4: return launch_fwd_ft<ST, ACC64, CPLX, 4>(st, g, tl, xv, out, os_b, br, bi);
        case 2: return launch_fwd_ft<ST, ACC64, CPLX, 2>(st, g, tl, xv, out, os_b, br, bi);
        default: return launch_fwd_ft<ST, ACC64, CPLX, 1>(st, g, tl, xv, out, os_b, br, bi);
    }
}

template <Storage ST, bool ACC64, bool CPLX, int FT>
hipError_t launch_bwdp_ft(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl,
                          const XView &xv, const void *gout, int64_t gs_b, float *gofs,
                          float *gwt, float *gscale) {
    constexpr bool PRECISE = (ST == Storage::F32 || ST == Storage::F64);
    const int ntiles = tl.nta * tl.nth * tl.ntw;
    dim3 grid((ntiles + 3) / 4, (g.B + FT - 1) / FT);
    dev::das_bwd_param_kernel<ST, ACC64, CPLX, PRECISE, FT, kKC><<<grid, 256, 0, st>>>(
        g, tl, xv.ptr, xv.sb, xv.sc, xv.st, gout, gs_b, gofs, gwt, gscale, grid.y > 1 ? 1 : 0);
    return hipGetLastError();
}

template <Storage ST, bool ACC64, bool CPLX>
hipError_t launch_bwdp_t(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl,
                         const XView &xv, const void *gout, int64_t gs_b, float *gofs,
                         float *gwt, float *gscale) {
    switch (frame_tile_for(g.B)) {
        case 16: return launch_bwdp_ft<ST, ACC64, CPLX, 16>(st, g, tl, xv, gout, gs_b, gofs, gwt, gscale);
        case 8: return launch_bwdp_ft<ST, ACC64, CPLX, 8>(st, g, tl, xv, gout, gs_b, gofs, gwt, gscale);
        case 4: return launch_bwdp_ft<ST, ACC64, CPLX, 4>(st, g, tl, xv, gout, gs_b, gofs, gwt, gscale);
        case 2: return launch_bwdp_ft<ST, ACC64, CPLX, 2>(st, g, tl, xv, gout, gs_b, gofs, gwt, gscale);
        default: return launch_bwdp_ft<ST, ACC64, CPLX, 1>(st, g, tl, xv, gout, gs_b, gofs, gwt, gscale);
    }
}

template <bool ACC64, bool CPLX, bool PRECISE>
hipError_t launch_bwdx_t(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl,
                         const void *gout, int64_t gs_b, void *gx, int FS) {
    const size_t esz = (ACC64 ? 8 : 4) * (CPLX ? 2 : 1);
    const size_t lds = (size_t)g.S * (size_t)(FS | 1) * esz;
    dim3 grid(g.C * g.T, (g.B + FS - 1) / FS);
    const int64_t gx_t = g.S, gx_c = (int64_t)g.T * g.S, gx_b = (int64_t)g.C * g.T * g.S;
    hipFuncSetAttribute((const void *)dev::das_bwd_x_kernel<ACC64, CPLX, PRECISE>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    dev::das_bwd_x_kernel<ACC64, CPLX, PRECISE><<<grid, 512, lds, st>>>(
        g, tl, gout, gs_b, gx, gx_b, gx_c, gx_t, FS);
    return hipGetLastError();
}

}  // namespace

ffdas_error_t launch_forward(ffdas_context &h, const DasProblem &p, const DasGeomArgs &a,
                             const XView &xv, void *out, double beta_re, double beta_im) {
    const dev::Geo g = make_geo(a, p);
    const dev::Tiling tl = make_tiling(p.spatial);
    hipError_t e = hipSuccess;
    hipStream_t st = h.stream;
    const int64_t os_b = p.os_b;
    if (!p.cplx) {
        switch (p.k_storage) {
            case Storage::F16: e = launch_fwd_t<Storage::F16, false, false>(st, g, tl, xv, out, os_b, beta_re, beta_im); break;
            case Storage::BF16: e = launch_fwd_t<Storage::BF16, false, false>(st, g, tl, xv, out, os_b, beta_re, beta_im); break;
            case Storage::F32:
                e = p.acc64 ? launch_fwd_t<Storage::F32, true, false>(st, g, tl, xv, out, os_b, beta_re, beta_im)
                            : launch_fwd_t<Storage::F32, false, false>(st, g, tl, xv, out, os_b, beta_re, beta_im);
                break;
            case Storage::F64: e = launch_fwd_t<Storage::F64, true, false>(st, g, tl, xv, out, os_b, beta_re, beta_im); break;
        }
    } else {
        switch (p.k_storage) {
            case Storage::F16: e = launch_fwd_t<Storage::F16, false, true>(st, g, tl, xv, out, os_b, beta_re, beta_im); break;
            case Storage::BF16: e = launch_fwd_t<Storage::BF16, false, true>(st, g, tl, xv, out, os_b, beta_re, beta_im); break;
            case Storage::F32:
                e = p.acc64 ? launch_fwd_t<Storage::F32, true, true>(st, g, tl, xv, out, os_b, beta_re, beta_im)
                            : launch_fwd_t<Storage::F32, false, true>(st, g, tl, xv, out, os_b, beta_re, beta_im);
                break;
            case Storage::F64: e = launch_fwd_t<Storage::F64, true, true>(st, g, tl, xv, out, os_b, beta_re, beta_im); break;
        }
    }
    return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

ffdas_error_t launch_fill_zero(ffdas_context &h, float *p, int64_t n) {
    if (!p || n <= 0) return FFDAS_SUCCESS;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    dev::fill_zero_kernel<<<(unsigned)blocks, 256, 0, h.stream>>>(p, n);
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

ffdas_error_t launch_backward(ffdas_context &h, const DasProblem &p, const DasGeomArgs &a,
                              const XView &xv, const void *gout, int64_t gs_b, void *gx,
                              float *gofs, float *gwt, float *gscale) {
    const dev::Geo g = make_geo(a, p);
    const dev::Tiling tl = make_tiling(p.spatial);
    hipStream_t st = h.stream;
    hipError_t e = hipSuccess;
    const bool precise = p.k_storage == Storage::F32 || p.k_storage == Storage::F64;

    if (gofs || gwt || gscale) {
        const int FT = frame_tile_for(p.B);
        const bool multi = (p.B + FT - 1) / FT > 1;
        if (multi) {
            FFDAS_TRY(launch_fill_zero(h, gofs, (int64_t)p.K * p.N));
            FFDAS_TRY(launch_fill_zero(h, gwt, (int64_t)p.K * p.N));
        }
        FFDAS_TRY(launch_fill_zero(h, gscale, 1));
#define BWDP(ST, A64, CX) e = launch_bwdp_t<ST, A64, CX>(st, g, tl, xv, gout, gs_b, gofs, gwt, gscale)
        if (!p.cplx) {
            switch (p.k_storage) {
                case Storage::F16: BWDP(Storage::F16, false, false); break;
                case Storage::BF16: BWDP(Storage::BF16, false, false); break;
                case Storage::F32: if (p.acc64) BWDP(Storage::F32, true, false); else BWDP(Storage::F32, false, false); break;
                case Storage::F64: BWDP(Storage::F64, true, false); break;
            }
        } else {
            switch (p.k_storage) {
                case Storage::F16: BWDP(Storage::F16, false, true); break;
                case Storage::BF16: BWDP(Storage::BF16, false, true); break;
                case Storage::F32: if (p.acc64) BWDP(Storage::F32, true, true); else BWDP(Storage::F32, false, true); break;
                case Storage::F64: BWDP(Storage::F64, true, true); break;
            }
        }
#undef BWDP
        if (e != hipSuccess) return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
    }
    if (gx) {
        const size_t esz = (p.acc64 ? 8 : 4) * (p.cplx ? 2 : 1);
        int FS = std::min(p.B, 16);
        while (FS > 0 && (size_t)p.S * (size_t)(FS | 1) * esz > (size_t)kLdsLimit) FS--;
        if (FS <= 0) return FFDAS_ERROR_DIMS_TOO_LARGE;
        if (p.acc64) {
            e = p.cplx ? launch_bwdx_t<true, true, true>(st, g, tl, gout, gs_b, gx, FS)
                       : launch_bwdx_t<true, false, true>(st, g, tl, gout, gs_b, gx, FS);
        } else if (precise) {
            e = p.cplx ? launch_bwdx_t<false, true, true>(st, g, tl, gout, gs_b, gx, FS)
                       : launch_bwdx_t<false, false, true>(st, g, tl, gout, gs_b, gx, FS);
        } else {
            e = p.cplx ? launch_bwdx_t<false, true, false>(st, g, tl, gout, gs_b, gx, FS)
                       : launch_bwdx_t<false, false, false>(st, g, tl, gout, gs_b, gx, FS);
        }
        if (e != hipSuccess) return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
    }
    return FFDAS_SUCCESS;
}

// Cast / relayout x into a contiguous (B, C, T, S) buffer of the kernel storage type.
ffdas_error_t launch_cast(ffdas_context &h, const DasProblem &p, const void *x, void *dst) {
    const int cw = p.cplx ? 2 : 1;
    const int64_t rows = (int64_t)p.B * p.C * p.T;
    if (rows > 0x7fffffff) return FFDAS_ERROR_DIMS_TOO_LARGE;
    dim3 grid((unsigned)rows);
    const int threads = 256;
    hipStream_t st = h.stream;
#define CAST(SRC, DST) \
    dev::cast_rows_kernel<SRC, DST><<<grid, threads, 0, st>>>((const SRC *)x, p.xs_b, p.xs_c, p.xs_t, (DST *)dst, p.C, p.T, p.S, cw)
    using bf = uint16_t;
    switch (p.x_storage) {
        case Storage::F32:
            switch (p.k_storage) {
                case Storage::F16: CAST(float, _Float16); break;
                case Storage::BF16: CAST(float, bf); break;
                case Storage::F32: CAST(float, float); break;
                default: return FFDAS_ERROR_INTERNAL;
            }
            break;
        case Storage::F64:
            switch (p.k_storage) {
                case Storage::F32: CAST(double, float); break;
                case Storage::F64: CAST(double, double); break;
                default: return FFDAS_ERROR_INTERNAL;
            }
            break;
        case Storage::F16:
            if (p.k_storage != Storage::F16) return FFDAS_ERROR_INTERNAL;
            CAST(_Float16, _Float16);
            break;
        case Storage::BF16:
            if (p.k_storage != Storage::BF16) return FFDAS_ERROR_INTERNAL;
            CAST(bf, bf);
            break;
    }
#undef CAST
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

}  // namespace ffdas
