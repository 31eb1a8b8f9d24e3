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
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <utility>

#include "das_device.h"

namespace ffdas {
namespace dev {

template <bool ACC64> using AccT = typename std::conditional<ACC64, double, float>::type;

__device__ __forceinline__ float fmaA(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaA(double a, double b, double c) { return __builtin_fma(a, b, c); }

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

// ---------------------------------------------------------------------------------------
// channel-data access policies
// ---------------------------------------------------------------------------------------
// Natural layout ([B,]C,T,S with S contiguous): FT frames per lane, one pointer per frame,
// x[k] and x[k+1] of each frame read separately.
template <Storage ST, int FT, int CW>
struct NaturalRows {
    using raw = typename Elem<ST>::raw;
    const raw *fb[FT];
    int64_t sc, st;
    __device__ __forceinline__ NaturalRows(const void *xv, int64_t xs_b, int64_t xs_c,
                                           int64_t xs_t, int b0, int B) {
        const raw *x = static_cast<const raw *>(xv);
#pragma unroll
        for (int f = 0; f < FT; f++) {
            const int b = min(b0 + f, B - 1);   // frames past B re-read the last one
            fb[f] = x + (int64_t)b * xs_b * CW;
        }
        sc = xs_c;
        st = xs_t;
    }
    __device__ __forceinline__ uint32_t offset(int m, int s, int k) const {
        return (uint32_t)((int64_t)m * sc) + (uint32_t)((int64_t)s * st) + (uint32_t)k;
    }
    // v[0][f][c] = x[k], v[1][f][c] = x[k+1]
    template <class A>
    __device__ __forceinline__ void load(uint32_t off, A (&v)[2][FT][CW]) const {
#pragma unroll
        for (int f = 0; f < FT; f++) {
            const raw *p = fb[f] + (size_t)off * CW;
#pragma unroll
            for (int c = 0; c < CW; c++) {
                v[0][f][c] = (A)Elem<ST>::real(p[c]);
                v[1][f][c] = (A)Elem<ST>::real(p[CW + c]);
            }
        }
    }
};

// Frame-inner layout, frames in groups of GS (written by the relayout kernels): per chunk
// [C][T][NG][S][GS][CW], NG = FT / GS, RB = GS * CW * esz bytes per row of a group.  The
// rows k and k+1 of a group are 2 * RB contiguous bytes, so a tap is NG * (2 * RB / 16)
// 16-byte loads, the groups S * RB bytes apart.  Two layouts:
// - planes (PL, dense problems): RB = 8 (4 FP16 or 2 FP32 frames per group).  The lanes of
//   a wave (neighbouring voxels, close sample indices, the same (channel, transmit) row)
//   then share cache lines, 16 samples of a plane per 128-byte line, which keeps the vector
//   L1's accesses per load low;
// - rows (!PL, sparse compounding): GS = FT, all frames of a row contiguous.  Neighbouring
//   voxels compound different transmits at a slot, so the lanes share few lines either way,
//   and one line per lane and tap instead of NG keeps the HBM traffic down.
// The rows of one channel are read through a buffer resource spanning that channel's block
// (below 2^31 bytes, which the host checks): the group step is a scalar offset, and a tap
// that reads nothing takes an offset past the block, for which the hardware returns zeros,
// so the loads of a group of taps need no branch and no select of addresses.
constexpr int kBufWord3 = 0x00020000;          // raw buffer, 32-bit data format
constexpr uint32_t kNoRows = 0x80000000u;      // beyond every channel block

template <int FT, int CW, int ESZ, bool PL>
constexpr int fi_group() { return PL ? 8 / (CW * ESZ) : FT; }

template <Storage ST, int FT, int CW, bool PL>
struct FrameInnerRows {
    using raw = typename Elem<ST>::raw;
    static constexpr int GS = fi_group<FT, CW, (int)sizeof(raw), PL>();   // frames per group
    static constexpr int NG = FT / GS;                          // groups per (channel, transmit)
    static constexpr int RB = GS * CW * (int)sizeof(raw);       // bytes of one row of a group
    static constexpr int LP = 2 * RB / 16;                      // 16-byte loads per group
    static constexpr int NQ = NG * LP;                          // 16-byte loads per tap
    static constexpr int E = 16 / (int)sizeof(raw);             // elements per load
    static_assert(GS >= 1 && FT % GS == 0 && (2 * RB) % 16 == 0, "frame groups must tile");
    struct Ref {
        __amdgpu_buffer_rsrc_t rows;   // the channel's block
        uint32_t off;                  // byte offset of rows k, k+1 of group 0
    };
    const raw *base;
    int T, S;
    __device__ __forceinline__ FrameInnerRows(const void *xv, int chunk, int C, int T_, int S_)
        : T(T_), S(S_) {
        base = static_cast<const raw *>(xv) + (int64_t)chunk * C * T_ * S_ * FT * CW;
    }
    __device__ __forceinline__ Ref offset(int m, int s, int k) const {
        const int mu = __builtin_amdgcn_readfirstlane(m);       // one channel per wave
        const raw *blk = base + (size_t)mu * T * S * FT * CW;
        Ref r;
        r.rows = __builtin_amdgcn_make_buffer_rsrc(const_cast<raw *>(blk), (short)0,
                                                   T * S * FT * CW * (int)sizeof(raw), kBufWord3);
        r.off = ((uint32_t)s * NG * (uint32_t)S + (uint32_t)k) * (uint32_t)RB;
        return r;
    }
    __device__ __forceinline__ void fetch(const Ref &r, uint4 (&q)[NQ]) const {
        fetch_at(r.rows, r.off, q);
    }
    // rows k, k+1 of (m, s), or zeros when the tap reads nothing
    __device__ __forceinline__ void fetch_or_zero(bool in, const Ref &r, uint4 (&q)[NQ]) const {
        fetch_at(r.rows, in ? r.off : kNoRows, q);
    }
    __device__ __forceinline__ void fetch_at(__amdgpu_buffer_rsrc_t rows, uint32_t off,
                                             uint4 (&q)[NQ]) const {
#pragma unroll
        for (int g = 0; g < NG; g++)
#pragma unroll
            for (int l = 0; l < LP; l++) {
                const auto v = __builtin_amdgcn_raw_buffer_load_b128(rows, (int)(off + 16 * l),
                                                                      g * S * RB, 0);
                q[g * LP + l] = make_uint4(v[0], v[1], v[2], v[3]);
            }
    }
    // v[h][f][c]: row k + h, frame f = g * GS + i, part c
    template <class A>
    __device__ __forceinline__ void unpack(const uint4 (&q)[NQ], A (&v)[2][FT][CW]) const {
#pragma unroll
        for (int g = 0; g < NG; g++) {
            raw r[2 * GS * CW];
            __builtin_memcpy(r, &q[g * LP], 2 * RB);
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int i = 0; i < GS; i++)
#pragma unroll
                    for (int c = 0; c < CW; c++)
                        v[h][g * GS + i][c] = (A)Elem<ST>::real(r[(h * GS + i) * CW + c]);
        }
    }
    template <class A>
    __device__ __forceinline__ void load(const Ref &r, A (&v)[2][FT][CW]) const {
        uint4 q[NQ];
        fetch(r, q);
        unpack(q, v);
    }
};

// ---------------------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------------------
template <class Rows, Storage ST, bool ACC64, bool CPLX, bool PRECISE, int FT, int KC>
__device__ __forceinline__ void das_fwd_body(const Geo &g, const Rows &xr, int n, int m0,
                                             int mstep, AccT<ACC64> (&acc)[FT][CPLX ? 2 : 1]) {
    using A = AccT<ACC64>;
    constexpr int CW = CPLX ? 2 : 1;
    const SplitPt yp = split_point(g.dstpos[3 * n], g.dstpos[3 * n + 1], g.dstpos[3 * n + 2]);
#pragma unroll
    for (int f = 0; f < FT; f++)
#pragma unroll
        for (int c = 0; c < CW; c++) acc[f][c] = A(0);

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
        for (int m = m0; m < g.C; m += mstep) {
            const SplitPt pp = split_point(g.srcpos[3 * m], g.srcpos[3 * m + 1], g.srcpos[3 * m + 2]);
            const Radius R = PRECISE ? radius_precise(yp, pp) : radius_fast(yp, pp);
            bool dok = true;
            if (g.srcdir) {
                const float *sd = g.srcdir + 4 * m;
                dok = (R.dx * sd[0] + R.dy * sd[1] + R.dz * sd[2]) * R.rinv > sd[3];
            }
#pragma unroll
            for (int j = 0; j < KC; j++) {
                if (kc + j >= g.K) break;
                const Tap tp = make_tap<PRECISE>(ofs[j], R, g.scale);
                const bool in = dok && (uint32_t)tp.k < (uint32_t)(g.S - 1) &&
                                (uint32_t)sq[j] < (uint32_t)g.T;
                if (in) {
                    A v[2][FT][CW];
                    xr.load(xr.offset(m, sq[j], tp.k), v);
                    if constexpr (!CPLX) {
                        const A c = (A)wt[j] * (A)tp.t;
                        const A a = (A)wt[j] - c;
#pragma unroll
                        for (int f = 0; f < FT; f++)
                            acc[f][0] = fmaA(c, v[1][f][0], fmaA(a, v[0][f][0], acc[f][0]));
                    } else {
                        float sn, cs;
                        phase<PRECISE>(g.rev_h, g.rev_l, tp, sn, cs);
                        const A wr = (A)(wt[j] * cs), wi = (A)(wt[j] * sn);
                        const A t = (A)tp.t;
#pragma unroll
                        for (int f = 0; f < FT; f++) {
                            const A lr = v[0][f][0] + t * (v[1][f][0] - v[0][f][0]);
                            const A li = v[0][f][1] + t * (v[1][f][1] - v[0][f][1]);
                            acc[f][0] += wr * lr - wi * li;
                            acc[f][1] += wr * li + wi * lr;
                        }
                    }
                }
            }
        }
    }
}

// Sparse compounding: the slots of a voxel in ascending transmit order (an odd-even merge
// sorting network on (index, offset, weight), 19 compare-exchanges for 8 slots).  The sum
// over slots does not depend on their order, but the lanes of a wave do: neighbouring
// voxels that compound the same transmits, ranked by distance in different orders, then
// read the same (channel, transmit) row at the same slot and share its cache lines.
template <int KC>
__device__ __forceinline__ void sort_slots(int (&sq)[KC], float (&ofs)[KC], float (&wt)[KC]) {
    auto cx = [&](int a, int b) {
        const bool sw = sq[a] > sq[b];
        const int s0 = sq[a], s1 = sq[b];
        const float o0 = ofs[a], o1 = ofs[b], w0 = wt[a], w1 = wt[b];
        sq[a] = sw ? s1 : s0; sq[b] = sw ? s0 : s1;
        ofs[a] = sw ? o1 : o0; ofs[b] = sw ? o0 : o1;
        wt[a] = sw ? w1 : w0; wt[b] = sw ? w0 : w1;
    };
    if constexpr (KC == 8) {
        cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
        cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
        cx(1, 2); cx(5, 6);
        cx(0, 4); cx(1, 5); cx(2, 6); cx(3, 7);
        cx(2, 4); cx(3, 5);
        cx(1, 2); cx(3, 4); cx(5, 6);
    } else {
#pragma unroll
        for (int i = 0; i < KC; i++)           // odd-even transposition sort
#pragma unroll
            for (int j = i & 1; j + 1 < KC; j += 2) cx(j, j + 1);
    }
}

// Frame-inner forward with the loads of TB taps in flight together: each group of TB taps
// issues all its 16-byte loads before the first of them is used, so a wave keeps TB times
// more gathers in flight than tap by tap.  A tap that reads nothing loads zeros (an offset
// past the channel's block) with its weights zeroed, so it adds exactly zero and no branch
// separates the loads from their use.
template <Storage ST, bool ACC64, bool CPLX, bool PRECISE, int FT, int KC, int TB, bool PL>
__device__ __forceinline__ void das_fwd_body_fi(const Geo &g, const FrameInnerRows<ST, FT, CPLX ? 2 : 1, PL> &xr,
                                                int n, int m0, int mstep,
                                                AccT<ACC64> (&acc)[FT][CPLX ? 2 : 1]) {
    static_assert(KC % TB == 0, "taps per group must divide taps per round");
    using A = AccT<ACC64>;
    using Rows = FrameInnerRows<ST, FT, CPLX ? 2 : 1, PL>;
    constexpr int CW = CPLX ? 2 : 1;
    const SplitPt yp = split_point(g.dstpos[3 * n], g.dstpos[3 * n + 1], g.dstpos[3 * n + 2]);
#pragma unroll
    for (int f = 0; f < FT; f++)
#pragma unroll
        for (int c = 0; c < CW; c++) acc[f][c] = A(0);
    if (g.S < 2) return;                       // no tap has two rows to read

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
                sq[j] = 0x7fffffff;       // selects nothing, sorts after every real slot
            }
        }
        if (g.sparse) sort_slots<KC>(sq, ofs, wt);
        for (int m = m0; m < g.C; m += mstep) {
            const SplitPt pp = split_point(g.srcpos[3 * m], g.srcpos[3 * m + 1], g.srcpos[3 * m + 2]);
            const Radius R = PRECISE ? radius_precise(yp, pp) : radius_fast(yp, pp);
            bool dok = true;
            if (g.srcdir) {
                const float *sd = g.srcdir + 4 * m;
                dok = (R.dx * sd[0] + R.dy * sd[1] + R.dz * sd[2]) * R.rinv > sd[3];
            }
#pragma unroll
            for (int j0 = 0; j0 < KC; j0 += TB) {
                if (kc + j0 >= g.K) break;
                uint4 q[TB][Rows::NQ];
                Tap tp[TB];
                bool in[TB];
#pragma unroll
                for (int u = 0; u < TB; u++) {
                    const int j = j0 + u;
                    tp[u] = make_tap<PRECISE>(ofs[j], R, g.scale);
                    in[u] = dok && (uint32_t)tp[u].k < (uint32_t)(g.S - 1) &&
                            (uint32_t)sq[j] < (uint32_t)g.T;
                    xr.fetch_or_zero(in[u], xr.offset(m, sq[j], tp[u].k), q[u]);
                }
                // keep the group's loads ahead of their first use
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < TB; u++) {
                    const int j = j0 + u;
                    A v[2][FT][CW];
                    xr.unpack(q[u], v);
                    if constexpr (!CPLX) {
                        const A c0 = (A)wt[j] * (A)tp[u].t;
                        const A c = in[u] ? c0 : A(0);
                        const A a = in[u] ? (A)wt[j] - c0 : A(0);
#pragma unroll
                        for (int f = 0; f < FT; f++)
                            acc[f][0] = fmaA(c, v[1][f][0], fmaA(a, v[0][f][0], acc[f][0]));
                    } else {
                        float sn, cs;
                        phase<PRECISE>(g.rev_h, g.rev_l, tp[u], sn, cs);
                        const A wr = in[u] ? (A)(wt[j] * cs) : A(0);
                        const A wi = in[u] ? (A)(wt[j] * sn) : A(0);
                        const A t = in[u] ? (A)tp[u].t : A(0);
#pragma unroll
                        for (int f = 0; f < FT; f++) {
                            const A lr = v[0][f][0] + t * (v[1][f][0] - v[0][f][0]);
                            const A li = v[0][f][1] + t * (v[1][f][1] - v[0][f][1]);
                            acc[f][0] += wr * lr - wi * li;
                            acc[f][1] += wr * li + wi * lr;
                        }
                    }
                }
            }
        }
    }
}

template <bool ACC64, bool CPLX, int FT>
__device__ __forceinline__ void store_out(const AccT<ACC64> (&acc)[FT][CPLX ? 2 : 1], void *outv,
                                          int64_t os_b, int n, int b0, int B, double beta_re,
                                          double beta_im) {
    using A = AccT<ACC64>;
    constexpr int CW = CPLX ? 2 : 1;
    A *out = static_cast<A *>(outv);
    const bool zero_beta = beta_re == 0.0 && beta_im == 0.0;
#pragma unroll
    for (int f = 0; f < FT; f++) {
        const int b = b0 + f;
        if (b >= B) break;
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

// natural layout (plain kernel; any strides)
template <Storage ST, bool ACC64, bool CPLX, bool PRECISE, int FT, int KC>
__global__ __launch_bounds__(256) void das_fwd_kernel(Geo g, Tiling tl, const void *__restrict__ xv,
                                                      int64_t xs_b, int64_t xs_c, int64_t xs_t,
                                                      void *__restrict__ outv, int64_t os_b,
                                                      double beta_re, double beta_im) {
    constexpr int CW = CPLX ? 2 : 1;
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    bool valid;
    const int n = voxel_of(tl, tile, lane, valid);
    if (!valid) return;
    const int b0 = blockIdx.y * FT;
    const NaturalRows<ST, FT, CW> xr(xv, xs_b, xs_c, xs_t, b0, g.B);
    AccT<ACC64> acc[FT][CW];
    das_fwd_body<NaturalRows<ST, FT, CW>, ST, ACC64, CPLX, PRECISE, FT, KC>(g, xr, n, 0, 1, acc);
    store_out<ACC64, CPLX, FT>(acc, outv, os_b, n, b0, g.B, beta_re, beta_im);
}

// frame-inner layout.  A workgroup of G waves shares one voxel tile; wave w sums channels
// w, w+G, ... and the G partial sums are combined through LDS, so a launch has G times more
// waves in flight than tiles (latency hiding for the gathers) without repeating any delay
// arithmetic.
template <Storage ST, bool ACC64, bool CPLX, bool PRECISE, int FT, int KC, int TB, bool PL>
__device__ __forceinline__ void fwd_fi_tile(const Geo &g, const Tiling &tl, const void *xv,
                                            void *outv, int64_t os_b, double beta_re,
                                            double beta_im, int tile, int chunk,
                                            unsigned char *lds) {
    using A = AccT<ACC64>;
    constexpr int CW = CPLX ? 2 : 1;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int G = blockDim.x >> 6;
    bool valid;
    int n = voxel_of(tl, tile, lane, valid);
    if (!valid) n = 0;          // computed, never stored: keeps every lane at the barrier
    const FrameInnerRows<ST, FT, CW, PL> xr(xv, chunk, g.C, g.T, g.S);
    A acc[FT][CW];
    if constexpr (TB == 1)
        das_fwd_body<FrameInnerRows<ST, FT, CW, PL>, ST, ACC64, CPLX, PRECISE, FT, KC>(g, xr, n, wave, G, acc);
    else
        das_fwd_body_fi<ST, ACC64, CPLX, PRECISE, FT, KC, TB, PL>(g, xr, n, wave, G, acc);
    if (G == 1) {
        if (valid) store_out<ACC64, CPLX, FT>(acc, outv, os_b, n, chunk * FT, g.B, beta_re, beta_im);
        return;
    }
    // red[w][i][lane], i = f * CW + c
    A *red = reinterpret_cast<A *>(lds);
#pragma unroll
    for (int f = 0; f < FT; f++)
#pragma unroll
        for (int c = 0; c < CW; c++) red[(wave * FT * CW + f * CW + c) * 64 + lane] = acc[f][c];
    __syncthreads();
    if (!valid) return;
    A *out = static_cast<A *>(outv);
    const bool zero_beta = beta_re == 0.0 && beta_im == 0.0;
    for (int f = wave; f < FT; f += G) {
        const int b = chunk * FT + f;
        if (b >= g.B) break;
        A v[CW];
#pragma unroll
        for (int c = 0; c < CW; c++) {
            A sum = A(0);
            for (int w = 0; w < G; w++) sum += red[(w * FT * CW + f * CW + c) * 64 + lane];
            v[c] = sum;
        }
        A *o = out + ((int64_t)b * os_b + n) * CW;
        if constexpr (!CPLX) {
            o[0] = zero_beta ? v[0] : (A)beta_re * o[0] + v[0];
        } else if (zero_beta) {
            o[0] = v[0];
            o[1] = v[1];
        } else {
            const A orr = o[0], oi = o[1];
            o[0] = (A)beta_re * orr - (A)beta_im * oi + v[0];
            o[1] = (A)beta_re * oi + (A)beta_im * orr + v[1];
        }
    }
}

// grid.x = voxel tile, grid.y = frame chunk
template <Storage ST, bool ACC64, bool CPLX, bool PRECISE, int FT, int KC, int TB, bool PL>
__global__ __launch_bounds__(512) void das_fwd_fi_kernel(Geo g, Tiling tl, const void *__restrict__ xv,
                                                         void *__restrict__ outv, int64_t os_b,
                                                         double beta_re, double beta_im) {
    extern __shared__ unsigned char fwd_lds[];
    int tile = blockIdx.x, chunk = blockIdx.y;
    if (tl.xcd) {
        const int b = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
        chunk = b / gridDim.x;
        tile = b - chunk * gridDim.x;
    }
    fwd_fi_tile<ST, ACC64, CPLX, PRECISE, FT, KC, TB, PL>(g, tl, xv, outv, os_b, beta_re, beta_im,
                                                       tile, chunk, fwd_lds);
}

// Several problems (mixed probes, grids, frame counts, sparse tables) in one launch.  The
// problem descriptors travel in the kernel arguments, so a captured graph holds them by
// value.  Block b belongs to the problem whose [start, start + tiles * chunks) contains it.
constexpr int kMaxBatch = 12;
struct BatchItem {
    Geo g;
    Tiling tl;
    const void *x;        // frame-inner copy of the problem's x
    void *out;
    int64_t os_b;
    double beta_re, beta_im;
    int start;            // first block of the problem
    int tiles;            // voxel tiles per frame chunk
};
struct BatchArgs {
    int count;
    int xcd;              // regroup blocks by XCD (xcd_block)
    BatchItem item[kMaxBatch];
};

template <Storage ST, bool CPLX, bool PRECISE, int FT, int KC, int TB, bool PL>
__global__ __launch_bounds__(512) void das_fwd_fi_batched_kernel(BatchArgs args) {
    extern __shared__ unsigned char fwd_lds[];
    const int b = args.xcd ? xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    int i = 0;
    while (i + 1 < args.count && b >= args.item[i + 1].start) i++;
    const BatchItem &it = args.item[i];
    const int local = b - it.start;
    const int chunk = local / it.tiles;
    fwd_fi_tile<ST, false, CPLX, PRECISE, FT, KC, TB, PL>(it.g, it.tl, it.x, it.out, it.os_b, it.beta_re,
                                                   it.beta_im, local - chunk * it.tiles, chunk,
                                                   fwd_lds);
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

template <class Rows, Storage ST, bool ACC64, bool CPLX, bool PRECISE, int FT, int KC>
__device__ __forceinline__ AccT<ACC64> das_bwd_param_body(const Geo &g, const Rows &xr, int n,
                                                          int b0, const void *goutv, int64_t gs_b,
                                                          float *gofs, float *gwt,
                                                          int m0, int mstep, bool valid,
                                                          AccT<ACC64> *red) {
    using A = AccT<ACC64>;
    constexpr int CW = CPLX ? 2 : 1;
    const A *__restrict__ gout = static_cast<const A *>(goutv);
    A scale_part = A(0);
    const SplitPt yp = split_point(g.dstpos[3 * n], g.dstpos[3 * n + 1], g.dstpos[3 * n + 2]);
    A gv[FT][CW];
#pragma unroll
    for (int f = 0; f < FT; f++) {
        const int b = b0 + f;
#pragma unroll
        for (int c = 0; c < CW; c++)
            gv[f][c] = (b < g.B) ? gout[((int64_t)b * gs_b + n) * CW + c] : A(0);
    }

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
        for (int m = m0; m < g.C; m += mstep) {
            const SplitPt pp = split_point(g.srcpos[3 * m], g.srcpos[3 * m + 1], g.srcpos[3 * m + 2]);
            const Radius R = PRECISE ? radius_precise(yp, pp) : radius_fast(yp, pp);
            bool dok = true;
            if (g.srcdir) {
                const float *sd = g.srcdir + 4 * m;
                dok = (R.dx * sd[0] + R.dy * sd[1] + R.dz * sd[2]) * R.rinv > sd[3];
            }
#pragma unroll
            for (int j = 0; j < KC; j++) {
                if (kc + j >= g.K) break;
                const Tap tp = make_tap<PRECISE>(ofs[j], R, g.scale);
                const bool in = dok && (uint32_t)tp.k < (uint32_t)(g.S - 1) &&
                                (uint32_t)sq[j] < (uint32_t)g.T;
                if (in) {
                    A v[2][FT][CW];
                    xr.load(xr.offset(m, sq[j], tp.k), v);
                    const A t = (A)tp.t;
                    // By linearity in the samples: with A0 = sum_f g x[k], A1 = sum_f g x[k+1],
                    // sum_f g (x[k+1] - x[k]) = A1 - A0 and sum_f g lerp = A0 + t (A1 - A0):
                    // two FMAs per frame instead of four.
                    if constexpr (!CPLX) {
                        A a0 = A(0), a1 = A(0);
#pragma unroll
                        for (int f = 0; f < FT; f++) {
                            a0 = fmaA(gv[f][0], v[0][f][0], a0);
                            a1 = fmaA(gv[f][0], v[1][f][0], a1);
                        }
                        const A sd_ = a1 - a0;
                        const A sl_ = fmaA(t, sd_, a0);
                        const A dt = (A)wt[j] * sd_;
                        d_tau[j] += dt;
                        d_w[j] += sl_;
                        scale_part += dt * ((A)ofs[j] + (A)R.rh + (A)R.rl);
                    } else {
                        // B0 = sum conj(g) x[k], B1 = sum conj(g) x[k+1];
                        // Q = B1 - B0 = sum conj(g) (x1 - x0), L = B0 + t Q = sum conj(g) lerp
                        A b0r = 0, b0i = 0, b1r = 0, b1i = 0;
#pragma unroll
                        for (int f = 0; f < FT; f++) {
                            const A gr = gv[f][0], gi = gv[f][1];
                            b0r = fmaA(gr, v[0][f][0], fmaA(gi, v[0][f][1], b0r));
                            b0i = fmaA(gr, v[0][f][1], fmaA(-gi, v[0][f][0], b0i));
                            b1r = fmaA(gr, v[1][f][0], fmaA(gi, v[1][f][1], b1r));
                            b1i = fmaA(gr, v[1][f][1], fmaA(-gi, v[1][f][0], b1i));
                        }
                        const A qr = b1r - b0r, qi = b1i - b0i;
                        const A lr = fmaA(t, qr, b0r), li = fmaA(t, qi, b0i);
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
        if (mstep > 1) {
            // combine the channel groups' partial sums: red[w][i][lane], i = 2 j + {0: tau, 1: w}
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
            for (int j = 0; j < KC; j++) {
                red[(wave * 2 * KC + 2 * j) * 64 + lane] = d_tau[j];
                red[(wave * 2 * KC + 2 * j + 1) * 64 + lane] = d_w[j];
            }
            __syncthreads();
            for (int i = wave; i < 2 * KC; i += mstep) {
                const int jj = kc + (i >> 1);
                A sum = A(0);
                for (int w = 0; w < mstep; w++) sum += red[(w * 2 * KC + i) * 64 + lane];
                float *dst = (i & 1) ? gwt : gofs;
                if (!valid || jj >= g.K || !dst) continue;
                const int64_t o = (int64_t)jj * g.N + n;
                dst[o] = (i & 1) ? (float)sum : (float)(sum * (A)g.scale);
            }
            __syncthreads();
            continue;
        }
#pragma unroll
        for (int j = 0; j < KC; j++) {
            const int jj = kc + j;
            if (jj >= g.K || !valid) break;
            const int64_t o = (int64_t)jj * g.N + n;
            if (gofs) gofs[o] = (float)(d_tau[j] * (A)g.scale);
            if (gwt) gwt[o] = (float)d_w[j];
        }
    }
    return scale_part;
}

// The block's part of d/d scale, stored at its block index (summed later in a fixed order,
// so the result does not depend on block scheduling).
template <bool ACC64>
__device__ __forceinline__ void block_scale_store(AccT<ACC64> part, double *partials) {
    using A = AccT<ACC64>;
    __shared__ A red[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    A v = wave_sum<ACC64>(part);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += (double)red[w];
        partials[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = s;
    }
}

// dst[i] = sum over c of part[c * n + i], c in order (frame-chunk partial gradients).
__global__ void sum_chunks_kernel(const float *__restrict__ part, int chunks, int64_t n,
                                  float *__restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int c = 0; c < chunks; c++) s += part[(int64_t)c * n + i];
        dst[i] = s;
    }
}

// *out = sum of n doubles in a fixed order (one block of 256 threads).
__global__ __launch_bounds__(256) void sum_partials_kernel(const double *__restrict__ part, int n,
                                                           float *__restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)red[0];
}

template <Storage ST, bool ACC64, bool CPLX, bool PRECISE, int FT, int KC>
__global__ __launch_bounds__(256) void das_bwd_param_kernel(
    Geo g, Tiling tl, const void *__restrict__ xv, int64_t xs_b, int64_t xs_c, int64_t xs_t,
    const void *__restrict__ goutv, int64_t gs_b, float *__restrict__ gofs,
    float *__restrict__ gwt, int64_t pstride, double *__restrict__ scale_part) {
    constexpr int CW = CPLX ? 2 : 1;
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    bool valid;
    const int n = voxel_of(tl, tile, lane, valid);
    AccT<ACC64> part = 0;
    if (valid) {
        const int b0 = blockIdx.y * FT;
        const NaturalRows<ST, FT, CW> xr(xv, xs_b, xs_c, xs_t, b0, g.B);
        // one frame chunk per grid row; its partial gradients go to their own slice
        float *go = gofs ? gofs + blockIdx.y * pstride : nullptr;
        float *gw = gwt ? gwt + blockIdx.y * pstride : nullptr;
        part = das_bwd_param_body<NaturalRows<ST, FT, CW>, ST, ACC64, CPLX, PRECISE, FT, KC>(
            g, xr, n, b0, goutv, gs_b, go, gw, 0, 1, true, nullptr);
    }
    if (scale_part) block_scale_store<ACC64>(part, scale_part);
}

// Frame-inner layout, G waves per voxel tile splitting the channels (as the forward).
template <Storage ST, bool ACC64, bool CPLX, bool PRECISE, int FT, int KC, bool PL>
__global__ __launch_bounds__(512) void das_bwd_param_fi_kernel(
    Geo g, Tiling tl, const void *__restrict__ xv, const void *__restrict__ goutv, int64_t gs_b,
    float *__restrict__ gofs, float *__restrict__ gwt, int64_t pstride,
    double *__restrict__ scale_part) {
    constexpr int CW = CPLX ? 2 : 1;
    extern __shared__ unsigned char param_lds[];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int G = blockDim.x >> 6;
    int tile = blockIdx.x, chunk = blockIdx.y;
    if (tl.xcd) {
        const int b = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
        chunk = b / gridDim.x;
        tile = b - chunk * gridDim.x;
    }
    bool valid;
    int n = voxel_of(tl, tile, lane, valid);
    if (!valid) n = 0;          // computed, never stored: keeps every lane at the barriers
    const FrameInnerRows<ST, FT, CW, PL> xr(xv, chunk, g.C, g.T, g.S);
    float *go = gofs ? gofs + chunk * pstride : nullptr;
    float *gw = gwt ? gwt + chunk * pstride : nullptr;
    AccT<ACC64> part = das_bwd_param_body<FrameInnerRows<ST, FT, CW, PL>, ST, ACC64, CPLX, PRECISE, FT, KC>(
        g, xr, n, chunk * FT, goutv, gs_b, go, gw, wave, G, valid,
        reinterpret_cast<AccT<ACC64> *>(param_lds));
    if (!valid) part = 0;
    if (scale_part) block_scale_store<ACC64>(part, scale_part);
}

// Relayout + cast into the frame-inner layout: dst[chunk][c][t][g][k][i] =
// src[chunk*FT + g*GS + i][c][t][k] (zero for frames past B), GS frames per group as read by
// FrameInnerRows.  One block handles a 64-sample segment of one (chunk, c, t) row for all FT
// frames, transposing through LDS so reads and writes are both contiguous.
template <Storage SS, class DST, int FT, int CW, int GS>
__global__ __launch_bounds__(256) void relayout_fi_kernel(const typename Elem<SS>::raw *__restrict__ src, int64_t sb,
                                                          int64_t sc, int64_t st, int B, int C,
                                                          int T, int S, DST *__restrict__ dst) {
    __shared__ float tile[FT][64 * CW + 1];
    const int nseg = (S + 63) >> 6;
    const int row = blockIdx.x / nseg;          // c * T + t (grid.x: rows x segments, so
    const int seg = blockIdx.x - row * nseg;    // C*T is not bound by the 65535 of grid.y)
    const int chunk = blockIdx.y;
    const int c = row / T, t = row - c * T;
    const int k0 = seg * 64;
    for (int i = threadIdx.x; i < FT * 64 * CW; i += blockDim.x) {
        const int f = i / (64 * CW), e = i - f * 64 * CW;
        const int k = k0 + e / CW;
        const int b = chunk * FT + f;
        float v = 0.0f;
        if (b < B && k < S) {
            const typename Elem<SS>::raw *p =
                src + ((int64_t)b * sb + (int64_t)c * sc + (int64_t)t * st) * CW;
            v = (float)Elem<SS>::real(p[(int64_t)k * CW + (e % CW)]);
        }
        tile[f][e] = v;
    }
    __syncthreads();
    DST *d = dst + ((((int64_t)chunk * C + c) * T + t) * S) * FT * CW;
    for (int i = threadIdx.x; i < 64 * FT * CW; i += blockDim.x) {
        // i = ((g * 64 + kk) * GS + fi) * CW + cc: consecutive threads write consecutive elements
        const int cc = i % CW, r = i / CW;
        const int fi = r % GS, r2 = r / GS;
        const int kk = r2 % 64, g = r2 / 64;
        const int k = k0 + kk;
        if (k < S)
            d[(((int64_t)g * S + k) * GS + fi) * CW + cc] = convert_to<DST>(tile[g * GS + fi][kk * CW + cc]);
    }
}

// The same relayout with 16-byte vector loads and stores, for rows whose start and length are
// 16-byte multiples: a block moves a segment of kRelL samples of one (chunk, c, t) row for
// all FT frames (each frame's segment read as 16-byte vectors, each plane's segment -- two
// samples of GS frames per vector -- written as 16-byte vectors), transposing through LDS.
constexpr int kRelL = 128;

template <Storage SS, class DST, int FT, int CW, int GS>
__global__ __launch_bounds__(256) void relayout_fi_vec_kernel(
    const typename Elem<SS>::raw *__restrict__ src, int64_t sb, int64_t sc, int64_t st, int B,
    int C, int T, int S, DST *__restrict__ dst) {
    using SR = typename Elem<SS>::raw;
    constexpr int SV = 16 / (int)sizeof(SR);        // source elements per vector
    constexpr int DV = 16 / (int)sizeof(DST);       // destination elements per vector
    constexpr int RL = kRelL * CW;                  // elements of one frame's segment
    __shared__ float tile[FT][RL + 1];
    const int nseg = (S + kRelL - 1) / kRelL;
    const int row = blockIdx.x / nseg;
    const int seg = blockIdx.x - row * nseg;
    const int chunk = blockIdx.y;
    const int c = row / T, t = row - c * T;
    const int e0 = seg * RL;                        // first element of the segment
    const int SE = S * CW;                          // elements per row
    for (int i = threadIdx.x; i < FT * (RL / SV); i += blockDim.x) {
        const int f = i / (RL / SV), e = (i - f * (RL / SV)) * SV;
        const int b = chunk * FT + f;
        float v[SV];
        if (b < B && e0 + e < SE) {
            const SR *p = src + ((int64_t)b * sb + (int64_t)c * sc + (int64_t)t * st) * CW + e0 + e;
            SR r[SV];
            const uint4 q = *reinterpret_cast<const uint4 *>(p);
            __builtin_memcpy(r, &q, 16);
#pragma unroll
            for (int j = 0; j < SV; j++) v[j] = (float)Elem<SS>::real(r[j]);
        } else {
#pragma unroll
            for (int j = 0; j < SV; j++) v[j] = 0.0f;
        }
#pragma unroll
        for (int j = 0; j < SV; j++) tile[f][e + j] = v[j];
    }
    __syncthreads();
    // each group's segment, kmax samples of GS * CW elements, is contiguous: written as
    // 16-byte vectors (the caller ensures kmax * GS * CW is a multiple of DV)
    DST *d = dst + (((int64_t)chunk * C + c) * T + t) * S * FT * CW;
    const int k0 = seg * kRelL;
    const int kmax = min(kRelL, S - k0);            // samples of this segment inside the row
    const int per_group = kmax * GS * CW / DV;      // vectors per group
    for (int i = threadIdx.x; i < FT / GS * per_group; i += blockDim.x) {
        const int g = i / per_group, e0 = (i - g * per_group) * DV;
        DST o[DV];
#pragma unroll
        for (int j = 0; j < DV; j++) {
            const int e = e0 + j;
            const int kk = e / (GS * CW), r = e - kk * GS * CW;
            const int fi = r / CW, cc = r - fi * CW;
            o[j] = convert_to<DST>(tile[g * GS + fi][kk * CW + cc]);
        }
        uint4 q;
        __builtin_memcpy(&q, o, 16);
        *reinterpret_cast<uint4 *>(d + ((int64_t)g * S + k0) * GS * CW + e0) = q;
    }
}

// ---------------------------------------------------------------------------------------
// adjoint: x (channel-row scatter through LDS)
// ---------------------------------------------------------------------------------------
// grad_out transposed to frame-inner chunks in axial voxel order:
// gt[chunk][q][f] = grad_out[chunk*FS+f][n(q)] (zero past B).  Chunk 0 also writes the
// geometry in axial order: pos[q], and ofs / wts / sp [j][q] for every slot j.  The largest
// |weight| and |grad_out| are folded into maxes[0] and maxes[1] (as float bits: non-negative
// floats order like their bit patterns), which must be zero on entry.
template <bool ACC64, bool CPLX, int FS>
__global__ __launch_bounds__(256) void transpose_gout_kernel(
    const void *__restrict__ goutv, int64_t gs_b, int B, int N, Axial ax,
    AccT<ACC64> *__restrict__ gt, Geo g, float *__restrict__ perm, unsigned *__restrict__ maxes) {
    using A = AccT<ACC64>;
    constexpr int CW = CPLX ? 2 : 1;
    const A *__restrict__ gout = static_cast<const A *>(goutv);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    float gmax = 0.0f, wmax = 0.0f;
    if (q < N) {
        const int r = q / ax.D0, d0 = q - r * ax.D0;
        const int n = d0 * ax.R + r;
        A v[FS * CW];
#pragma unroll
        for (int f = 0; f < FS; f++) {
            const int b = chunk * FS + f;
#pragma unroll
            for (int c = 0; c < CW; c++) {
                v[f * CW + c] = b < B ? gout[((int64_t)b * gs_b + n) * CW + c] : A(0);
                gmax = fmaxf(gmax, (float)fabs((double)v[f * CW + c]));
            }
        }
        A *dst = gt + ((int64_t)chunk * N + q) * FS * CW;
#pragma unroll
        for (int i = 0; i < FS * CW; i++) dst[i] = v[i];
        if (chunk == 0) {
            float *pos = perm;
            float *ofs = pos + 3 * (int64_t)N;
            float *wts = ofs + (int64_t)g.K * N;
            int *sp = reinterpret_cast<int *>(wts + (int64_t)g.K * N);
            pos[3 * (int64_t)q] = g.dstpos[3 * (int64_t)n];
            pos[3 * (int64_t)q + 1] = g.dstpos[3 * (int64_t)n + 1];
            pos[3 * (int64_t)q + 2] = g.dstpos[3 * (int64_t)n + 2];
            // eight rows loaded before any is stored: a store may alias the next load as far as
            // the compiler knows, so a row-by-row loop waited a memory round trip per row
            for (int j0 = 0; j0 < g.K; j0 += 8) {
                float ov[8], wv[8];
                int sv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int j = j0 + u;
                    const bool in = j < g.K;
                    ov[u] = in ? g.offsets[(int64_t)j * N + n] : 0.0f;
                    wv[u] = in ? g.weights[(int64_t)j * N + n] : 0.0f;
                    sv[u] = in && g.sparse ? g.sparse[(int64_t)j * N + n] : 0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int j = j0 + u;
                    if (j >= g.K) break;
                    ofs[(int64_t)j * N + q] = ov[u];
                    wts[(int64_t)j * N + q] = wv[u];
                    wmax = fmaxf(wmax, fabsf(wv[u]));
                    if (g.sparse) sp[(int64_t)j * N + q] = sv[u];
                }
            }
        }
    }
    // block maximum first: one atomic per block and maximum
    __shared__ float red[2][4];
    for (int o = 32; o > 0; o >>= 1) {
        gmax = fmaxf(gmax, __shfl_xor(gmax, o, 64));
        wmax = fmaxf(wmax, __shfl_xor(wmax, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = gmax;
        red[1][threadIdx.x >> 6] = wmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            gmax = fmaxf(gmax, red[0][w]);
            wmax = fmaxf(wmax, red[1][w]);
        }
        if (gmax > 0.0f) atomicMax(maxes + 1, __float_as_uint(gmax));
        if (chunk == 0 && wmax > 0.0f) atomicMax(maxes, __float_as_uint(wmax));
    }
}

// Fixed-point accumulation: the contribution w * g (g already scaled by 2^E) is rounded to
// an integer by one FP64 FMA with the magic constant 1.5 * 2^52 (the integer then sits in the
// low mantissa bits, exact for |w g 2^E| < 2^51) and added to the 64-bit LDS accumulator with
// an integer atomic: sums are exact and order-independent (LDS integer atomics run ~40x faster
// than LDS float atomics on gfx950).
constexpr double kFxMagic = 6755399441055744.0;                 // 1.5 * 2^52
constexpr unsigned long long kFxMagicBits = 0x4338000000000000ull;

__device__ __forceinline__ void fx_add(unsigned long long *p, double w, double g_scaled) {
    const double y = __builtin_fma(w, g_scaled, kFxMagic);
    atomicAdd(p, (unsigned long long)__double_as_longlong(y) - kFxMagicBits);
}

// round(w1 g1 + round(w2 g2)): two products into one accumulator (complex parts)
__device__ __forceinline__ void fx_add2(unsigned long long *p, double w1, double g1, double w2,
                                        double g2) {
    const double y = __builtin_fma(w1, g1, __builtin_fma(w2, g2, kFxMagic));
    atomicAdd(p, (unsigned long long)__double_as_longlong(y) - kFxMagicBits);
}

// Inverted index of a sparse table (K, N) of transmit indices: for every transmit s the
// entries q * K + j (voxel q in axial order, slot j) with sparse[j][n(q)] == s, so the grad_x
// row of transmit s visits only them.  Entries with an index outside [0, T) select nothing.
// Counts and placements are aggregated per block in LDS (one global atomic per block and
// transmit); the order inside a transmit's range is not fixed, the fixed-point sums it feeds
// do not depend on it.
constexpr int kMaxSparseT = 4096;

__global__ __launch_bounds__(256) void sparse_count_kernel(const int *__restrict__ sparse,
                                                           int64_t entries, int T,
                                                           unsigned *__restrict__ count) {
    __shared__ unsigned h[kMaxSparseT];
    const bool local = T <= kMaxSparseT;
    if (local)
        for (int i = threadIdx.x; i < T; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < entries;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int t = sparse[e];
        if ((unsigned)t >= (unsigned)T) continue;
        if (local) atomicAdd(h + t, 1u);
        else atomicAdd(count + t, 1u);
    }
    __syncthreads();
    if (local)
        for (int i = threadIdx.x; i < T; i += blockDim.x)
            if (h[i]) atomicAdd(count + i, h[i]);
}

// start[t] = sum of count[0..t), start[T] = total (one thread; T is a transmit count).
__global__ void sparse_scan_kernel(const unsigned *__restrict__ count, int T,
                                   unsigned *__restrict__ start) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    unsigned acc = 0;
    for (int t = 0; t < T; t++) {
        start[t] = acc;
        acc += count[t];
    }
    start[T] = acc;
}

__global__ __launch_bounds__(256) void sparse_fill_kernel(const int *__restrict__ sparse, int K,
                                                          int N, int T, Axial ax,
                                                          const unsigned *__restrict__ start,
                                                          unsigned *__restrict__ cursor,
                                                          int *__restrict__ list) {
    __shared__ unsigned h[kMaxSparseT], base[kMaxSparseT];
    const bool local = T <= kMaxSparseT;
    const int64_t entries = (int64_t)K * N;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (local)
        for (int i = threadIdx.x; i < T; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    int t = -1;
    unsigned rank = 0;
    if (e < entries) {
        t = sparse[e];
        if ((unsigned)t >= (unsigned)T) t = -1;
    }
    if (t >= 0 && local) rank = atomicAdd(h + t, 1u);
    __syncthreads();
    if (local)
        for (int i = threadIdx.x; i < T; i += blockDim.x)
            base[i] = h[i] ? atomicAdd(cursor + i, h[i]) : 0u;
    __syncthreads();
    if (t < 0) return;
    const unsigned pos = start[t] + (local ? base[t] + rank : atomicAdd(cursor + t, 1u));
    const int j = (int)(e / N), n = (int)(e - (int64_t)j * N);
    const int d0 = n / ax.R, r = n - d0 * ax.R;      // n = d0 * R + r  ->  q = r * D0 + d0
    list[pos] = (r * ax.D0 + d0) * K + j;
}

// Block-ordered inverted index (T <= kMaxSparseT): the axial voxels are cut into blocks of
// kSpBlock; sparse_bcount_kernel histograms each block's entries per transmit in LDS,
// sparse_bscan_kernel scans each transmit's block counts (and sparse_scan_kernel the
// transmit totals), and sparse_bfill_kernel writes every block's entries of transmit t
// contiguously at that block's offset.  Each transmit's list is thus ordered by voxel block
// (order inside a block is free), so a row of the scatter visits neighbouring voxels
// together, as the dense rows do.  sp is the sparse table in axial order, (K, N).
constexpr int kSpBlock = 1024;

__global__ __launch_bounds__(256) void sparse_bcount_kernel(const int *__restrict__ sp, int K,
                                                            int N, int T, int nblk,
                                                            unsigned *__restrict__ tc) {
    __shared__ unsigned h[kMaxSparseT];
    for (int i = threadIdx.x; i < T; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    const int q0 = blockIdx.x * kSpBlock;
    const int q1 = min(N, q0 + kSpBlock);
    for (int q = q0 + (int)threadIdx.x; q < q1; q += blockDim.x)
        for (int j = 0; j < K; j++) {
            const int t = sp[(int64_t)j * N + q];
            if ((unsigned)t < (unsigned)T) atomicAdd(h + t, 1u);
        }
    __syncthreads();
    for (int i = threadIdx.x; i < T; i += blockDim.x) tc[(int64_t)i * nblk + blockIdx.x] = h[i];
}

// one block per transmit t: exclusive scan of tc[t][0..nblk) in place, tot[t] = the sum
__global__ __launch_bounds__(256) void sparse_bscan_kernel(unsigned *__restrict__ tc, int nblk,
                                                           unsigned *__restrict__ tot) {
    __shared__ unsigned part[256];
    unsigned *row = tc + (int64_t)blockIdx.x * nblk;
    unsigned carry = 0;
    for (int b0 = 0; b0 < nblk; b0 += 256) {
        const int b = b0 + (int)threadIdx.x;
        const unsigned v = b < nblk ? row[b] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {     // inclusive Hillis-Steele scan
            const unsigned add = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (b < nblk) row[b] = carry + part[threadIdx.x] - v;
        carry += part[255];
        __syncthreads();
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

__global__ __launch_bounds__(256) void sparse_bfill_kernel(const int *__restrict__ sp, int K,
                                                           int N, int T, int nblk,
                                                           const unsigned *__restrict__ tc,
                                                           const unsigned *__restrict__ start,
                                                           int *__restrict__ list) {
    __shared__ unsigned h[kMaxSparseT];
    for (int i = threadIdx.x; i < T; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    const int q0 = blockIdx.x * kSpBlock;
    const int q1 = min(N, q0 + kSpBlock);
    for (int q = q0 + (int)threadIdx.x; q < q1; q += blockDim.x)
        for (int j = 0; j < K; j++) {
            const int t = sp[(int64_t)j * N + q];
            if ((unsigned)t >= (unsigned)T) continue;
            const unsigned rank = atomicAdd(h + t, 1u);
            list[start[t] + tc[(int64_t)t * nblk + blockIdx.x] + rank] = q * K + j;
        }
}

// LDS slot of sample k: within each 32-sample block the low five bits are XORed with seven
// times the block number, so samples a multiple of 4 or 8 apart (neighbouring voxels along
// depth) land in different LDS banks.  A bijection on [0, 32 * ceil(S / 32)).  Measured on the
// benchmark (FP16 storage, 16 frames per workgroup): 0.549-0.552 ms against 0.555-0.560 ms
// for the block number itself, 0.703 ms without the XOR and 0.59-0.61 ms for a rotation by
// the block number; lanes spread over distant columns instead of neighbouring ones changed
// nothing at 16 lanes along depth.
__device__ __forceinline__ int lds_slot(int k) { return k ^ (((k >> 5) * 7) & 31); }

// The row is stored sample-major: the WS words of slot s (frame pairs times complex parts)
// at s * PS, PS = WS rounded up to odd, so that one instruction's lanes (all at the same
// word w) map slot -> bank pair (s * PS + w) mod 32 bijectively, and the words of one slot
// are immediate offsets from one address.
__host__ __device__ constexpr int bwdx_pitch(int ws) { return ws % 2 == 0 ? ws + 1 : ws; }

// Packed fixed point: two frames share one 64-bit LDS word, frame 2p in the low and frame
// 2p+1 in the high 32 bits, each an integer sum in units of 2^-E.  Every contribution c is
// below 2^22 in magnitude (the exponent guarantees it), so the FP32 FMA y = w g + 1.5*2^23
// forms w g exactly and rounds it to the integer c, which then sits in the mantissa:
// bits(y) = 0x4B400000 + c.  The addend is the pair of raw bits (bits(y0), bits(y1)) as one
// 64-bit integer; after n contributions the word holds
//   n * 0x4B400000 * (1 + 2^32) + sum0 + sum1 * 2^32   (mod 2^64),
// and with n from the per-sample count the bias is removed; as long as |sum0|, |sum1| < 2^31
// the low half sign-extended is sum0 and the rest is sum1.
constexpr float kFxMagicF = 12582912.0f;                        // 1.5 * 2^23
constexpr unsigned long long kFxMagicBitsF = 0x4B400000ull;

__device__ __forceinline__ void fx_add_pair(unsigned long long *p, float y0, float y1) {
    atomicAdd(p, ((unsigned long long)__float_as_uint(y1) << 32) | __float_as_uint(y0));
}

__device__ __forceinline__ void fx_unpack_pair(unsigned long long w, unsigned n, long long &s0,
                                               long long &s1) {
    const unsigned long long bias = (unsigned long long)n * kFxMagicBitsF;
    w -= bias + (bias << 32);
    s0 = (long long)(int)(unsigned)w;
    s1 = ((long long)w - s0) >> 32;
}

// Largest value in the block (all threads take part, all receive it).
__device__ __forceinline__ unsigned block_max_u32(unsigned v, unsigned *red) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned r = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) r = max(r, red[w]);
    __syncthreads();
    return r;
}

// grid.x = rows row0 .. row0 + gridDim.x - 1 of the C*T rows (m = row / T, s = row % T),
// grid.y = frame chunks of FS frames.  The transposed grad_out holds records of fsl frames
// per voxel (fsl a multiple of FS), so a launch with half the frames reads half records.
// The row lives in LDS as fixed point [slot(k)][word] (word = frame or frame pair, times the
// complex part, at an odd pitch), so the lanes of one atomic differ only in slot; each lane owns
// one voxel per step (axial order: geometry and grad_out were permuted to it, so the
// per-wave loads are contiguous) and adds its two interpolation weights times grad_out
// into the row.  The loads of a lane's next voxel are issued before the current voxel's
// atomics.
//
// PACK (float data): two frames per 64-bit word with 32-bit sums.  The scale 2^E is chosen
// per row from max|w| * max|g| and the largest number of contributions any sample of the
// row receives: a first pass assumes at most 2^kLogCnt0 and counts the contributions per
// sample in LDS; a row that exceeded it is accumulated again with the exponent its count
// allows.  No 32-bit sum can overflow, and the rounding error per contribution is below
// 2^-(E+1) (about 2^-24 of the largest contribution in the usual case).
// !PACK (double data, or one frame per chunk): one 64-bit integer per frame, with the
// global bound of N (or N*K sparse) contributions per sample.
constexpr int kLogCnt0 = 8;

template <bool ACC64, bool CPLX, bool PRECISE, int FS>
__global__ __launch_bounds__(1024) void das_bwd_x_kernel(Geo g, Axial ax, const void *__restrict__ gtv,
                                                         const unsigned *__restrict__ maxes,
                                                         int log2cnt, void *__restrict__ gxv,
                                                         int64_t gx_b, int64_t gx_c, int64_t gx_t,
                                                         const unsigned *__restrict__ sp_start,
                                                         const int *__restrict__ sp_list,
                                                         int row0, int fsl) {
    using A = AccT<ACC64>;
    using U64 = unsigned long long;
    constexpr bool PACK = !ACC64 && FS >= 2;
    constexpr int CW = CPLX ? 2 : 1;
    constexpr int NW = PACK ? FS / 2 : FS;                 // words per sample and part
    constexpr int WS = NW * CW;                            // words per sample
    constexpr int PS = bwdx_pitch(WS);                     // words between samples (odd)
    extern __shared__ __align__(16) unsigned char lds_raw[];
    U64 *lds = reinterpret_cast<U64 *>(lds_raw);
    const int S32 = (g.S + 31) & ~31;
    unsigned *cnt = reinterpret_cast<unsigned *>(lds + (size_t)S32 * PS);   // PACK only
    __shared__ unsigned red[16];
    const A *__restrict__ gt = static_cast<const A *>(gtv);
    A *__restrict__ gx = static_cast<A *>(gxv);

    const int row = row0 + (int)blockIdx.x;
    const int m = row / g.T;
    const int s = row - m * g.T;
    const int chunk = blockIdx.y;
    const int b0 = chunk * FS;
    const int nf = min(FS, g.B - b0);

    // exponent of max|w| * max|g| (bound < 2^eM)
    int eM = 0;
    {
        const double bound = (double)__uint_as_float(maxes[0]) * (double)__uint_as_float(maxes[1]);
        if (bound > 0.0) frexp(bound, &eM);
    }
    // PACK scales the weights in FP32: 2^E must keep max|w| * 2^E finite (a grad_out below
    // about 2^(-100) max|w|... is then resolved more coarsely than 2^-24 of its largest term)
    int eW = 0;
    if (PACK) {
        const float wmax = __uint_as_float(maxes[0]);
        if (wmax > 0.0f) frexpf(wmax, &eW);
    }
    auto exponent = [&](int logc) {
        // PACK: sums < 2^30 and contributions < 2^22 (FP32 trick range, logc >= 8);
        // !PACK: contributions < 2^min(40, 61 - logc) (FP64 trick range, 63-bit sums)
        const int top = PACK ? 30 - logc - (CPLX ? 1 : 0) : min(40, 61 - logc);
        int E = top - eM;
        if (PACK) E = max(-126, min(E, 126 - eW));
        return E < -900 ? -900 : (E > 900 ? 900 : E);
    };
    int E = exponent(PACK ? kLogCnt0 : log2cnt);

    const SplitPt pp = split_point(g.srcpos[3 * m], g.srcpos[3 * m + 1], g.srcpos[3 * m + 2]);
    float sdx = 0, sdy = 0, sdz = 0, sdw = 0;
    if (g.srcdir) {
        sdx = g.srcdir[4 * m]; sdy = g.srcdir[4 * m + 1]; sdz = g.srcdir[4 * m + 2];
        sdw = g.srcdir[4 * m + 3];
    }
    const bool dense = g.sparse == nullptr;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int waves = blockDim.x >> 6;
    const int D0T = 1 << ax.lg, RT = 64 >> ax.lg;
    const int dl = lane & (D0T - 1), rl = lane >> ax.lg;
    const int per_r = (ax.D0 + D0T - 1) >> ax.lg;
    // dense: every voxel of the grid; sparse: the (voxel, slot) entries of the inverted
    // index that select this row's transmit, 64 per wave step
    const unsigned e_lo = dense ? 0u : sp_start[s];
    const unsigned e_hi = dense ? 0u : sp_start[s + 1];
    const int nitems = dense ? per_r * ((ax.R + RT - 1) / RT) : (int)((e_hi - e_lo + 63u) >> 6);
    const int64_t rec = b0 / fsl;   // the record group holding frames b0 .. b0 + FS - 1
    const A *gchunk = gt + (rec * g.N * fsl + (b0 - rec * fsl)) * CW;

    struct Vox {
        A gv[FS * CW];
        float x, y, z, ofs, w;
        int q;
        bool valid;
    };
    auto fetch = [&](int item, Vox &v) {
        int64_t o = 0;    // offsets / weights entry of the voxel's slot
        if (dense) {
            const int rg = item / per_r;
            const int d0 = (item - rg * per_r) * D0T + dl;
            const int r = rg * RT + rl;
            v.valid = d0 < ax.D0 && r < ax.R;
            v.q = v.valid ? r * ax.D0 + d0 : 0;     // invalid lanes read voxel 0, never scatter
            o = (int64_t)s * g.N + v.q;
        } else {
            const unsigned e = e_lo + (unsigned)item * 64u + (unsigned)lane;
            v.valid = e < e_hi;
            const int code = v.valid ? sp_list[e] : 0;      // q * K + j
            v.q = code / g.K;
            o = (int64_t)(code - v.q * g.K) * g.N + v.q;
        }
        const A *gp = gchunk + (int64_t)v.q * fsl * CW;
#pragma unroll
        for (int i = 0; i < FS * CW; i++) v.gv[i] = gp[i];
        v.x = g.dstpos[3 * (int64_t)v.q];
        v.y = g.dstpos[3 * (int64_t)v.q + 1];
        v.z = g.dstpos[3 * (int64_t)v.q + 2];
        v.ofs = g.offsets[o];
        v.w = g.weights[o];
    };

    for (int pass = 0; pass < 2; pass++) {
        const bool counting = PACK && pass == 0;   // pass 1 keeps pass 0's counts
        const double sc = ldexp(1.0, E);
        const float scf = PACK ? ldexpf(1.0f, E) : 0.0f;
        for (int i = threadIdx.x; i < S32 * PS; i += blockDim.x) lds[i] = 0ull;
        if (counting)
            for (int i = threadIdx.x; i < S32; i += blockDim.x) cnt[i] = 0u;
        __syncthreads();

        auto scatter = [&](const Vox &v, const Radius &R, float ofs, float w) {
            const Tap tp = make_tap<PRECISE>(ofs, R, g.scale);
            if ((uint32_t)tp.k >= (uint32_t)(g.S - 1)) return;
            const int s0 = lds_slot(tp.k), s1 = lds_slot(tp.k + 1);
            U64 *l0 = lds + s0 * PS;
            U64 *l1 = lds + s1 * PS;
            // one count per tap, at its first sample: sample s then received
            // cnt(s) + cnt(s - 1) contributions
            if (counting) atomicAdd(cnt + s0, 1u);
            if constexpr (PACK) {
                // FP32: weights scaled by 2^E, the FMA rounds each product to its integer
                const float tf = tp.t;
                const float wf = w * scf;
                if constexpr (!CPLX) {
                    const float w1 = wf * tf, w0 = wf - w1;
#pragma unroll
                    for (int p = 0; p < NW; p++) {
                        const float ga = (float)v.gv[2 * p], gb = (float)v.gv[2 * p + 1];
                        fx_add_pair(l0 + p, __builtin_fmaf(w0, ga, kFxMagicF),
                                    __builtin_fmaf(w0, gb, kFxMagicF));
                        fx_add_pair(l1 + p, __builtin_fmaf(w1, ga, kFxMagicF),
                                    __builtin_fmaf(w1, gb, kFxMagicF));
                    }
                } else {
                    float sn, cs;
                    phase<PRECISE>(g.rev_h, g.rev_l, tp, sn, cs);
                    const float cr = wf * cs, ci = wf * sn;
                    const float omt = 1.0f - tf;
                    // conj(coef) * g = (cr gr + ci gi) + i (cr gi - ci gr), split by (1-t), t
                    const float a_r = cr * omt, a_i = ci * omt, b_r = cr * tf, b_i = ci * tf;
                    // words [p][re, im], each packing frames 2p, 2p+1
#pragma unroll
                    for (int p = 0; p < NW; p++) {
                        const float gr0 = (float)v.gv[4 * p], gi0 = (float)v.gv[4 * p + 1];
                        const float gr1 = (float)v.gv[4 * p + 2], gi1 = (float)v.gv[4 * p + 3];
                        fx_add_pair(l0 + 2 * p,
                                    __builtin_fmaf(a_r, gr0, __builtin_fmaf(a_i, gi0, kFxMagicF)),
                                    __builtin_fmaf(a_r, gr1, __builtin_fmaf(a_i, gi1, kFxMagicF)));
                        fx_add_pair(l0 + 2 * p + 1,
                                    __builtin_fmaf(a_r, gi0, __builtin_fmaf(-a_i, gr0, kFxMagicF)),
                                    __builtin_fmaf(a_r, gi1, __builtin_fmaf(-a_i, gr1, kFxMagicF)));
                        fx_add_pair(l1 + 2 * p,
                                    __builtin_fmaf(b_r, gr0, __builtin_fmaf(b_i, gi0, kFxMagicF)),
                                    __builtin_fmaf(b_r, gr1, __builtin_fmaf(b_i, gi1, kFxMagicF)));
                        fx_add_pair(l1 + 2 * p + 1,
                                    __builtin_fmaf(b_r, gi0, __builtin_fmaf(-b_i, gr0, kFxMagicF)),
                                    __builtin_fmaf(b_r, gi1, __builtin_fmaf(-b_i, gr1, kFxMagicF)));
                    }
                }
                return;
            } else {
                // one 64-bit integer per frame: the FP64 FMA rounds each product
                const double t = (double)tp.t;
                const double wd = (double)w * sc;
                if constexpr (!CPLX) {
                    const double w1 = wd * t, w0 = wd - w1;
#pragma unroll
                    for (int f = 0; f < FS; f++) {
                        const double gd = (double)v.gv[f];
                        fx_add(l0 + f, w0, gd);
                        fx_add(l1 + f, w1, gd);
                    }
                } else {
                    float sn, cs;
                    phase<PRECISE>(g.rev_h, g.rev_l, tp, sn, cs);
                    const double cr = wd * (double)cs, ci = wd * (double)sn;
                    const double omt = 1.0 - t;
                    const double a_r = cr * omt, a_i = ci * omt, b_r = cr * t, b_i = ci * t;
#pragma unroll
                    for (int f = 0; f < FS; f++) {
                        const double gr = (double)v.gv[2 * f], gi = (double)v.gv[2 * f + 1];
                        fx_add2(l0 + 2 * f, a_r, gr, a_i, gi);
                        fx_add2(l0 + 2 * f + 1, a_r, gi, -a_i, gr);
                        fx_add2(l1 + 2 * f, b_r, gr, b_i, gi);
                        fx_add2(l1 + 2 * f + 1, b_r, gi, -b_i, gr);
                    }
                }
            }
        };
        auto process = [&](const Vox &v) {
            if (!v.valid) return;
            const SplitPt yp = split_point(v.x, v.y, v.z);
            const Radius R = PRECISE ? radius_precise(yp, pp) : radius_fast(yp, pp);
            if (g.srcdir && !((R.dx * sdx + R.dy * sdy + R.dz * sdz) * R.rinv > sdw)) return;
            scatter(v, R, v.ofs, v.w);
        };

        // two voxel buffers in turn: the next voxel's loads overlap this one's scatter
        Vox va, vb;
        va.valid = false;
        vb.valid = false;
        int item = wave;
        if (item < nitems) fetch(item, va);
        while (item < nitems) {
            vb.valid = false;
            if (item + waves < nitems) fetch(item + waves, vb);
            process(va);
            item += waves;
            if (item >= nitems) break;
            va.valid = false;
            if (item + waves < nitems) fetch(item + waves, va);
            process(vb);
            item += waves;
        }
        __syncthreads();
        if (!counting) break;
        // the row's largest contribution count decides whether pass 0's sums are valid
        unsigned cmax = 0;
        for (int i = threadIdx.x; i < g.S; i += blockDim.x)
            cmax = max(cmax, cnt[lds_slot(i)] + (i > 0 ? cnt[lds_slot(i - 1)] : 0u));
        cmax = block_max_u32(cmax, red);
        if (cmax <= (1u << kLogCnt0)) break;
        int logc = kLogCnt0;
        while ((1ull << logc) < cmax) logc++;
        E = exponent(logc);
    }

    // write back: consecutive threads -> consecutive samples, streamed past the caches (grad_x
    // is written once and not read again here; a step that reuses its channel data then
    // still finds that data in the last-level cache)
    const double inv = ldexp(1.0, -E);
    for (int k = threadIdx.x; k < g.S; k += blockDim.x) {
        const U64 *wp = lds + (size_t)lds_slot(k) * PS;
        const unsigned nk = PACK ? cnt[lds_slot(k)] + (k > 0 ? cnt[lds_slot(k - 1)] : 0u) : 0u;
        A *dst = gx + ((int64_t)b0 * gx_b + (int64_t)m * gx_c + (int64_t)s * gx_t + k) * CW;
        if constexpr (PACK) {
#pragma unroll
            for (int p = 0; p < NW; p++) {
#pragma unroll
                for (int c = 0; c < CW; c++) {
                    long long v0, v1;
                    fx_unpack_pair(wp[p * CW + c], nk, v0, v1);
                    if (2 * p < nf)
                        __builtin_nontemporal_store((A)((double)v0 * inv),
                                                    dst + (int64_t)(2 * p) * gx_b * CW + c);
                    if (2 * p + 1 < nf)
                        __builtin_nontemporal_store((A)((double)v1 * inv),
                                                    dst + (int64_t)(2 * p + 1) * gx_b * CW + c);
                }
            }
        } else {
#pragma unroll
            for (int f = 0; f < FS; f++) {
                if (f >= nf) break;
#pragma unroll
                for (int c = 0; c < CW; c++)
                    __builtin_nontemporal_store((A)((double)(long long)wp[f * CW + c] * inv),
                                                dst + (int64_t)f * gx_b * CW + c);
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// adjoint: x, plain form (ALG1)
// ---------------------------------------------------------------------------------------
// The straightforward transpose of the plain forward kernel: one lane per voxel with the
// grad_out of FT frames in registers, every in-range (voxel, channel, slot) adds its two
// interpolation weights times grad_out into grad_x with global float atomics.  grad_x must
// be zero on entry.  Summation order is not fixed, so results vary in the last bits.
template <bool ACC64, bool CPLX, bool PRECISE, int FT, int KC>
__global__ __launch_bounds__(256) void das_bwd_x_plain_kernel(
    Geo g, Tiling tl, const void *__restrict__ goutv, int64_t gs_b, void *__restrict__ gxv,
    int64_t gx_b, int64_t gx_c, int64_t gx_t) {
    using A = AccT<ACC64>;
    constexpr int CW = CPLX ? 2 : 1;
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    bool valid;
    const int n = voxel_of(tl, tile, lane, valid);
    if (!valid) return;
    const A *__restrict__ gout = static_cast<const A *>(goutv);
    A *__restrict__ gx = static_cast<A *>(gxv);
    const int b0 = blockIdx.y * FT;
    const int nf = min(FT, g.B - b0);
    A gv[FT][CW];
#pragma unroll
    for (int f = 0; f < FT; f++)
#pragma unroll
        for (int c = 0; c < CW; c++)
            gv[f][c] = (b0 + f < g.B) ? gout[((int64_t)(b0 + f) * gs_b + n) * CW + c] : A(0);
    const SplitPt yp = split_point(g.dstpos[3 * n], g.dstpos[3 * n + 1], g.dstpos[3 * n + 2]);
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
            if (g.srcdir) {
                const float *sd = g.srcdir + 4 * m;
                if (!((R.dx * sd[0] + R.dy * sd[1] + R.dz * sd[2]) * R.rinv > sd[3])) continue;
            }
#pragma unroll
            for (int j = 0; j < KC; j++) {
                if (kc + j >= g.K) break;
                const Tap tp = make_tap<PRECISE>(ofs[j], R, g.scale);
                if (!((uint32_t)tp.k < (uint32_t)(g.S - 1) && (uint32_t)sq[j] < (uint32_t)g.T)) continue;
                A *row = gx + ((int64_t)b0 * gx_b + (int64_t)m * gx_c + (int64_t)sq[j] * gx_t +
                               tp.k) * CW;
                const A t = (A)tp.t;
                if constexpr (!CPLX) {
                    const A w1 = (A)wt[j] * t, w0 = (A)wt[j] - w1;
                    for (int f = 0; f < nf; f++) {
                        atomicAdd(row + f * gx_b, w0 * gv[f][0]);
                        atomicAdd(row + f * gx_b + 1, w1 * gv[f][0]);
                    }
                } else {
                    float sn, cs;
                    phase<PRECISE>(g.rev_h, g.rev_l, tp, sn, cs);
                    const A cr = (A)wt[j] * (A)cs, ci = (A)wt[j] * (A)sn;
                    for (int f = 0; f < nf; f++) {
                        // conj(w e^{i phi}) g
                        const A vr = cr * gv[f][0] + ci * gv[f][1];
                        const A vi = cr * gv[f][1] - ci * gv[f][0];
                        A *q = row + f * gx_b * CW;
                        atomicAdd(q, (A(1) - t) * vr);
                        atomicAdd(q + 1, (A(1) - t) * vi);
                        atomicAdd(q + 2, t * vr);
                        atomicAdd(q + 3, t * vi);
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// adjoint: x by bucketing voxels per sample (gather form of the scatter)
// ---------------------------------------------------------------------------------------
// One workgroup (1024 threads) owns one channel-data row (channel m, transmit s) for a chunk
// of F frames.  Voxels are processed in chunks of nc: every contributing (voxel, slot) is
// counted into its sample bucket k with an LDS integer atomic (whose return value is the
// record's slot in the bucket), the bucket starts come from a block-wide exclusive scan, and
// the records {voxel, (1-t) w, t w} are placed bucket-ordered in LDS.  Thread i owns samples
// [KPT*i, KPT*i + KPT) for all F frames and sums, in FP32 registers, w0 * g into sample k and
// w1 * g into sample k+1 for every record of its buckets: one FMA per contribution and no
// atomics on the data.  The sample just past a thread's range is handed to the next thread
// (shuffle within a wave, LDS across waves).  Summation order within a bucket follows the
// atomic ranks, so results can differ from run to run in the last bits; the fixed-point
// kernel above is the bitwise-deterministic alternative.
template <bool CPLX> struct BwdRec;
template <> struct BwdRec<false> { int n; float w0, w1; };
template <> struct BwdRec<true> { int n; float ar, ai, br, bi; };

template <int NT>
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned *wsum) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (wave == 0) {
        unsigned t = lane < NT / 64 ? wsum[lane] : 0u;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned y = __shfl_up(t, o, 64);
            if (lane >= o) t += y;
        }
        if (lane < NT / 64) wsum[lane] = t;   // inclusive wave totals
    }
    __syncthreads();
    const unsigned before = wave > 0 ? wsum[wave - 1] : 0u;
    return before + x - v;
}

constexpr int kSortNT = 1024;                   // threads per workgroup
template <bool CPLX> constexpr int sort_vpt() { return 2; }   // records per thread per chunk

// Records of bucket k0+A add (1-t) w g to acc[A] and t w g to acc[A+1]; A is a template
// parameter so every accumulator index is static (registers, not scratch).
template <int A, int KPT, int F, bool CPLX>
__device__ __forceinline__ void sorted_bucket_pass(float (&acc)[KPT + 1][F][CPLX ? 2 : 1],
                                                   const BwdRec<CPLX> *rec, const unsigned *cnt,
                                                   const float *gchunk, int k0, int S) {
    constexpr int CW = CPLX ? 2 : 1;
    const int k = k0 + A;
    if (k >= S) return;
    const unsigned r0 = cnt[k], r1 = cnt[k + 1];
    for (unsigned r = r0; r < r1; r++) {
        const BwdRec<CPLX> rc_ = rec[r];
        const float *gp = gchunk + (int64_t)rc_.n * F * CW;
        if constexpr (!CPLX) {
            constexpr int H = F > 8 ? 8 : F;      // frames per load group
#pragma unroll
            for (int h = 0; h < F; h += H) {
                float gv[H];
#pragma unroll
                for (int i = 0; i < H; i++) gv[i] = gp[h + i];
#pragma unroll
                for (int i = 0; i < H; i++) {
                    acc[A][h + i][0] = __builtin_fmaf(rc_.w0, gv[i], acc[A][h + i][0]);
                    acc[A + 1][h + i][0] = __builtin_fmaf(rc_.w1, gv[i], acc[A + 1][h + i][0]);
                }
            }
        } else {
            float gv[F * CW];
#pragma unroll
            for (int i = 0; i < F * CW; i++) gv[i] = gp[i];
#pragma unroll
            for (int f = 0; f < F; f++) {
                const float gr = gv[2 * f], gi = gv[2 * f + 1];
                // conj(a) g and conj(b) g
                acc[A][f][0] += rc_.ar * gr + rc_.ai * gi;
                acc[A][f][1] += rc_.ar * gi - rc_.ai * gr;
                acc[A + 1][f][0] += rc_.br * gr + rc_.bi * gi;
                acc[A + 1][f][1] += rc_.br * gi - rc_.bi * gr;
            }
        }
    }
    if constexpr (A + 1 < KPT) sorted_bucket_pass<A + 1, KPT, F, CPLX>(acc, rec, cnt, gchunk, k0, S);
}

template <bool CPLX, bool PRECISE, int KPT, int F>
__global__ __launch_bounds__(1024) void das_bwd_x_sorted_kernel(
    Geo g, const float *__restrict__ gtf, float *__restrict__ gx, int64_t gx_b, int64_t gx_c,
    int64_t gx_t, int nc) {
    constexpr int NT = kSortNT;
    constexpr int CW = CPLX ? 2 : 1;
    constexpr int VPT = sort_vpt<CPLX>();
    using Rec = BwdRec<CPLX>;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    unsigned *cnt = reinterpret_cast<unsigned *>(lds_raw);            // S+1 bucket starts
    unsigned *wsum = cnt + (g.S + 1);                                  // NT/64 scan partials
    float *xwave = reinterpret_cast<float *>(wsum + NT / 64);          // NT/64 * F * CW carries
    Rec *rec = reinterpret_cast<Rec *>(xwave + (NT / 64) * F * CW);

    const int row = blockIdx.x;
    const int m = row / g.T;
    const int s = row - m * g.T;
    const int chunk = blockIdx.y;
    const int b0 = chunk * F;
    const int nf = min(F, g.B - b0);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int k0 = tid * KPT;

    float acc[KPT + 1][F][CW];
#pragma unroll
    for (int a = 0; a <= KPT; a++)
#pragma unroll
        for (int f = 0; f < F; f++)
#pragma unroll
            for (int c = 0; c < CW; c++) acc[a][f][c] = 0.0f;

    const SplitPt pp = split_point(g.srcpos[3 * m], g.srcpos[3 * m + 1], g.srcpos[3 * m + 2]);
    float sdx = 0, sdy = 0, sdz = 0, sdw = 0;
    if (g.srcdir) {
        sdx = g.srcdir[4 * m]; sdy = g.srcdir[4 * m + 1]; sdz = g.srcdir[4 * m + 2];
        sdw = g.srcdir[4 * m + 3];
    }
    const int kslots = g.sparse ? g.K : 1;
    const float *gchunk = gtf + (int64_t)chunk * g.N * F * CW;

    for (int v0 = 0; v0 < g.N; v0 += nc) {
        for (int i = tid; i <= g.S; i += NT) cnt[i] = 0u;
        __syncthreads();
        // phase A: taps of this chunk's voxels -> bucket counts (+ rank within the bucket)
        int rk[VPT], rn[VPT];
        unsigned rr[VPT];
        float ra[VPT], rb[VPT], rc[VPT], rd[VPT];
#pragma unroll
        for (int q = 0; q < VPT; q++) {
            rk[q] = -1;
            const int li = q * NT + tid;            // candidate (voxel, slot) in the chunk
            const int vi = li / kslots;
            const int j = g.sparse ? li - vi * kslots : s;
            const int n = v0 + vi;
            if (vi >= nc || n >= g.N) continue;
            const int64_t o = (int64_t)j * g.N + n;
            if (g.sparse && g.sparse[o] != s) continue;
            const SplitPt yp = split_point(g.dstpos[3 * n], g.dstpos[3 * n + 1], g.dstpos[3 * n + 2]);
            const Radius R = PRECISE ? radius_precise(yp, pp) : radius_fast(yp, pp);
            if (g.srcdir && !((R.dx * sdx + R.dy * sdy + R.dz * sdz) * R.rinv > sdw)) continue;
            const Tap tp = make_tap<PRECISE>(g.offsets[o], R, g.scale);
            if ((uint32_t)tp.k >= (uint32_t)(g.S - 1)) continue;
            const float w = g.weights[o];
            rk[q] = tp.k;
            rn[q] = n;
            if constexpr (!CPLX) {
                rb[q] = w * tp.t;
                ra[q] = w - rb[q];
            } else {
                float sn, cs;
                phase<PRECISE>(g.rev_h, g.rev_l, tp, sn, cs);
                const float omt = 1.0f - tp.t;
                ra[q] = w * cs * omt; rb[q] = w * sn * omt;
                rc[q] = w * cs * tp.t; rd[q] = w * sn * tp.t;
            }
            rr[q] = atomicAdd(&cnt[tp.k], 1u);
        }
        __syncthreads();
        // exclusive scan of the S+1 counts, each thread over a contiguous run of <= 8
        {
            constexpr int maxrun = 8;
            const int run = (g.S + 1 + NT - 1) / NT;
            unsigned loc[maxrun];
            unsigned tot = 0;
#pragma unroll
            for (int r = 0; r < maxrun; r++) {
                const int i = tid * run + r;
                loc[r] = (r < run && i <= g.S) ? cnt[i] : 0u;
                tot += loc[r];
            }
            const unsigned base = block_exclusive_scan<NT>(tot, wsum);
            unsigned acc_s = base;
#pragma unroll
            for (int r = 0; r < maxrun; r++) {
                const int i = tid * run + r;
                if (r < run && i <= g.S) cnt[i] = acc_s;
                acc_s += loc[r];
            }
        }
        __syncthreads();
        // placement
#pragma unroll
        for (int q = 0; q < VPT; q++) {
            if (rk[q] < 0) continue;
            Rec &r = rec[cnt[rk[q]] + rr[q]];
            r.n = rn[q];
            if constexpr (!CPLX) {
                r.w0 = ra[q]; r.w1 = rb[q];
            } else {
                r.ar = ra[q]; r.ai = rb[q]; r.br = rc[q]; r.bi = rd[q];
            }
        }
        __syncthreads();
        // phase D: accumulate the records of the owned buckets
        sorted_bucket_pass<0, KPT, F, CPLX>(acc, rec, cnt, gchunk, k0, g.S);
        __syncthreads();
    }
    // hand the sample past each thread's range to the next thread, then write
#pragma unroll
    for (int f = 0; f < F; f++)
#pragma unroll
        for (int c = 0; c < CW; c++) {
            const float up = __shfl_up(acc[KPT][f][c], 1, 64);
            if (lane == 63) xwave[(wave * F + f) * CW + c] = acc[KPT][f][c];
            if (lane > 0) acc[0][f][c] += up;
        }
    __syncthreads();
    if (lane == 0 && wave > 0) {
#pragma unroll
        for (int f = 0; f < F; f++)
#pragma unroll
            for (int c = 0; c < CW; c++) acc[0][f][c] += xwave[((wave - 1) * F + f) * CW + c];
    }
#pragma unroll
    for (int f = 0; f < F; f++) {
        if (f >= nf) break;
        float *dst = gx + ((int64_t)(b0 + f) * gx_b + (int64_t)m * gx_c + (int64_t)s * gx_t) * CW;
#pragma unroll
        for (int a = 0; a < KPT; a++) {
            const int k = k0 + a;
            if (k >= g.S) break;
#pragma unroll
            for (int c = 0; c < CW; c++) dst[(size_t)k * CW + c] = acc[a][f][c];
        }
    }
}

// grad_out transposed to frame-inner float chunks: gt[chunk][n][f] (zero past B).
template <bool CPLX, int F>
__global__ void transpose_gout_f_kernel(const float *__restrict__ gout, int64_t gs_b, int B, int N,
                                        float *__restrict__ gt) {
    constexpr int CW = CPLX ? 2 : 1;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    if (n >= N) return;
    float v[F * CW];
#pragma unroll
    for (int f = 0; f < F; f++) {
        const int b = chunk * F + f;
#pragma unroll
        for (int c = 0; c < CW; c++)
            v[f * CW + c] = b < B ? gout[((int64_t)b * gs_b + n) * CW + c] : 0.0f;
    }
    float *dst = gt + ((int64_t)chunk * N + n) * F * CW;
#pragma unroll
    for (int i = 0; i < F * CW; i++) dst[i] = v[i];
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

// The gathers' workgroups regrouped so that each XCD takes a contiguous range of tiles
// (neighbouring voxel tiles, whose rows overlap, then share that XCD's L2).  Measured with
// the frame-group planes: 2-D benchmark forward 0.198 -> 0.192 ms, parameter adjoint
// 0.226 -> 0.223 ms; batched sparse compounding (mixed probes, HBM-bound) 575 -> 534 ms.
// FFDAS_XCD=0 keeps the hardware's round-robin order.
bool xcd_groups() {
    static const bool on = [] {
        const char *env = std::getenv("FFDAS_XCD");
        return !(env && std::atoi(env) == 0);
    }();
    return on;
}

dev::Tiling make_tiling(const std::vector<int64_t> &spatial, const DasTune &tune = DasTune()) {
    std::vector<int64_t> d;
    for (int64_t v : spatial)
        if (v > 1) d.push_back(v);
    dev::Tiling t{};
    t.A = 1; t.H = 1; t.W = 1;
    if (d.size() >= 1) t.W = (int)d.back();
    if (d.size() >= 2) t.H = (int)d[d.size() - 2];
    for (size_t i = 0; i + 2 < d.size(); i++) t.A *= (int)d[i];
    if (d.size() >= 3) { t.ta = 4; t.th = 4; t.tw = 4; }
    else if (d.size() == 2) {
        t.ta = 1; t.th = 8; t.tw = 8;
        if (tune.tile_h > 0 && tune.tile_w > 0 && tune.tile_h * tune.tile_w == 64) {
            t.th = tune.tile_h;
            t.tw = tune.tile_w;
        }
    }
    else { t.ta = 1; t.th = 1; t.tw = 64; }
    t.nta = (t.A + t.ta - 1) / t.ta;
    t.nth = (t.H + t.th - 1) / t.th;
    t.ntw = (t.W + t.tw - 1) / t.tw;
    t.xcd = xcd_groups() ? 1 : 0;
    return t;
}

dev::Axial make_axial(const DasProblem &p) {
    dev::Axial ax{};
    ax.D0 = 1;
    for (int64_t v : p.spatial)
        if (v > 1) { ax.D0 = (int)v; break; }
    ax.R = (int)(p.N / ax.D0);
    ax.lg = (p.tune.bwdx_lg >= 0 && p.tune.bwdx_lg <= 6) ? p.tune.bwdx_lg : 4;
    if (const char *env = std::getenv("FFDAS_BWDX_AXIAL")) ax.lg = std::max(0, std::min(6, std::atoi(env)));
    while (ax.lg > 0 && (1 << (ax.lg - 1)) >= ax.D0) ax.lg--;
    return ax;
}

// Waves per voxel tile for the frame-inner gathers: enough to put about 32 waves on each
// CU (8 per SIMD), at most 8 and at most one per channel.  FFDAS_CGROUPS overrides.
int channel_groups(const dev::Geo &g, int64_t tiles, int want) {
    int G = 1;
    if (const char *env = std::getenv("FFDAS_CGROUPS")) {
        G = std::atoi(env);
    } else if (want > 0) {
        G = want;
    } else {
        while (G < 8 && tiles * G < 8192) G *= 2;
    }
    G = std::max(1, std::min(G, 8));
    while (G > 1 && G > g.C) G >>= 1;
    return G;
}

dev::Geo make_geo(const DasGeomArgs &a, const DasProblem &p) {
    dev::Geo g{};
    g.srcpos = a.srcpos; g.srcdir = a.srcdir; g.dstpos = a.dstpos;
    g.offsets = a.offsets; g.weights = a.weights; g.sparse = p.sparse ? a.sparse_indices : nullptr;
    g.C = p.C; g.T = p.T; g.S = p.S; g.K = p.K; g.B = p.B; g.N = (int)p.N;
    g.scale = a.scale;
    g.wavenum = a.wavenum;
    const double rev = (double)a.wavenum / (2.0 * 3.14159265358979323846);
    g.rev_h = (float)rev;
    g.rev_l = (float)(rev - (double)g.rev_h);
    return g;
}

template <Storage ST, bool ACC64, bool CPLX, int FT>
hipError_t launch_fwd_ft(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl, const XView &xv,
                         void *out, int64_t os_b, double br, double bi) {
    constexpr bool PRECISE = true;
    const int ntiles = tl.nta * tl.nth * tl.ntw;
    dim3 grid((ntiles + 3) / 4, (g.B + FT - 1) / FT);
    ProfScope ps(st, "das_fwd");
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

// Final and partial buffers of the parameter adjoint: with several frame chunks each chunk
// writes its own slice of pofs / pwt, and every block its d/d scale part to spart; the
// sums are then formed in a fixed order.
struct ParamOut {
    float *gofs = nullptr, *gwt = nullptr, *gscale = nullptr;
    float *pofs = nullptr, *pwt = nullptr;
    double *spart = nullptr;
};

hipError_t finish_params(hipStream_t st, const ParamOut &po, int chunks, int64_t kn, int blocks) {
    if (chunks > 1) {
        const int grid = (int)std::min<int64_t>((kn + 255) / 256, 4096);
        if (po.gofs) dev::sum_chunks_kernel<<<grid, 256, 0, st>>>(po.pofs, chunks, kn, po.gofs);
        if (po.gwt) dev::sum_chunks_kernel<<<grid, 256, 0, st>>>(po.pwt, chunks, kn, po.gwt);
    }
    if (po.gscale) dev::sum_partials_kernel<<<1, 256, 0, st>>>(po.spart, blocks, po.gscale);
    return hipGetLastError();
}

template <Storage ST, bool ACC64, bool CPLX, int FT>
hipError_t launch_bwdp_ft(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl,
                          const XView &xv, const void *gout, int64_t gs_b, const ParamOut &po) {
    constexpr bool PRECISE = true;
    const int ntiles = tl.nta * tl.nth * tl.ntw;
    dim3 grid((ntiles + 3) / 4, (g.B + FT - 1) / FT);
    const bool multi = grid.y > 1;
    const int64_t kn = (int64_t)g.K * g.N;
    {
    ProfScope ps(st, "das_bwd_param");
    dev::das_bwd_param_kernel<ST, ACC64, CPLX, PRECISE, FT, kKC><<<grid, 256, 0, st>>>(
        g, tl, xv.ptr, xv.sb, xv.sc, xv.st, gout, gs_b, multi ? po.pofs : po.gofs,
        multi ? po.pwt : po.gwt, multi ? kn : 0, po.gscale ? po.spart : nullptr);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return finish_params(st, po, (int)grid.y, kn, (int)(grid.x * grid.y));
}

template <Storage ST, bool ACC64, bool CPLX>
hipError_t launch_bwdp_t(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl,
                         const XView &xv, const void *gout, int64_t gs_b, const ParamOut &po) {
    switch (frame_tile_for(g.B)) {
        case 16: return launch_bwdp_ft<ST, ACC64, CPLX, 16>(st, g, tl, xv, gout, gs_b, po);
        case 8: return launch_bwdp_ft<ST, ACC64, CPLX, 8>(st, g, tl, xv, gout, gs_b, po);
        case 4: return launch_bwdp_ft<ST, ACC64, CPLX, 4>(st, g, tl, xv, gout, gs_b, po);
        case 2: return launch_bwdp_ft<ST, ACC64, CPLX, 2>(st, g, tl, xv, gout, gs_b, po);
        default: return launch_bwdp_ft<ST, ACC64, CPLX, 1>(st, g, tl, xv, gout, gs_b, po);
    }
}

// LDS of the grad_x kernel: the fixed-point row (two frames per 64-bit word for float data,
// one for double data or a single frame) plus, packed, the per-sample contribution counts.
size_t bwdx_lds_bytes(int S, int FS, bool cplx, bool acc64) {
    const size_t S32 = (size_t)((S + 31) & ~31);
    const bool pack = !acc64 && FS >= 2;
    const int ws = (pack ? FS / 2 : FS) * (cplx ? 2 : 1);
    return S32 * (size_t)dev::bwdx_pitch(ws) * 8 + (pack ? S32 * 4 : 0);
}

template <bool ACC64, bool CPLX, bool PRECISE, int FT>
hipError_t launch_bwdx_plain_ft(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl,
                                const void *gout, int64_t gs_b, void *gx) {
    const int ntiles = tl.nta * tl.nth * tl.ntw;
    dim3 grid((ntiles + 3) / 4, (g.B + FT - 1) / FT);
    const int64_t gx_t = g.S, gx_c = (int64_t)g.T * g.S, gx_b = (int64_t)g.C * g.T * g.S;
    ProfScope ps(st, "das_bwd_x_plain");
    dev::das_bwd_x_plain_kernel<ACC64, CPLX, PRECISE, FT, kKC><<<grid, 256, 0, st>>>(
        g, tl, gout, gs_b, gx, gx_b, gx_c, gx_t);
    return hipGetLastError();
}

template <bool ACC64, bool CPLX, bool PRECISE>
hipError_t launch_bwdx_plain_t(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl,
                               const void *gout, int64_t gs_b, void *gx) {
    switch (frame_tile_for(g.B)) {
        case 16: return launch_bwdx_plain_ft<ACC64, CPLX, PRECISE, 16>(st, g, tl, gout, gs_b, gx);
        case 8: return launch_bwdx_plain_ft<ACC64, CPLX, PRECISE, 8>(st, g, tl, gout, gs_b, gx);
        case 4: return launch_bwdx_plain_ft<ACC64, CPLX, PRECISE, 4>(st, g, tl, gout, gs_b, gx);
        case 2: return launch_bwdx_plain_ft<ACC64, CPLX, PRECISE, 2>(st, g, tl, gout, gs_b, gx);
        default: return launch_bwdx_plain_ft<ACC64, CPLX, PRECISE, 1>(st, g, tl, gout, gs_b, gx);
    }
}

// Rows of a dense grad_x launch given to a second launch with half the frames per workgroup:
// with one row-chunk per workgroup, the last round of workgroups often fills only part of
// the CUs (1408 rows of 16 frames on 256 CUs: 5.5 rounds).  When that last round is at most
// half full, its rows go to workgroups of FS / 2 frames (two per row-chunk, each about 0.6
// of the time), and the first launch covers whole rounds.  The sums are the same integers
// either way (the exponent depends on neither FS nor the chunking), so grad_x is bitwise
// unchanged.  0 = no split.
int bwdx_tail_rows(int rows, int nchunks, int cu_count, int per_cu, int per_cu_half) {
    if (cu_count <= 0 || per_cu <= 0 || per_cu_half <= 0) return 0;
    const int64_t cap = (int64_t)cu_count * per_cu, cap2 = (int64_t)cu_count * per_cu_half;
    if ((int64_t)rows * nchunks <= cap) return 0;   // a single round: nothing to balance
    int tr = 0;
    while (tr < rows && ((int64_t)(rows - tr) * nchunks) % cap != 0) tr++;
    if (tr == 0 || tr == rows) return 0;
    const int64_t w2 = (int64_t)tr * nchunks * 2;
    return w2 <= cap && w2 <= cap2 ? tr : 0;
}

template <bool ACC64, bool CPLX, bool PRECISE, int FS>
hipError_t launch_bwdx_fs(hipStream_t st, const dev::Geo &g, const dev::Axial &ax, int threads,
                          const void *gout, int64_t gs_b, void *gt, void *gx, int cu_count) {
    using A = typename std::conditional<ACC64, double, float>::type;
    const size_t esz = 8 * (CPLX ? 2 : 1);
    const size_t lds = bwdx_lds_bytes(g.S, FS, CPLX, ACC64);
    const int nchunks = (g.B + FS - 1) / FS;
    // workspace: [maxes: 2 words, padded to 256 B][transposed grad_out][axial geometry]
    unsigned *maxes = static_cast<unsigned *>(gt);
    void *gtr = static_cast<char *>(gt) + 256;
    float *perm = reinterpret_cast<float *>(static_cast<char *>(gtr) +
                                            (size_t)nchunks * FS * g.N * esz);
    // contributions per LDS sample at most N (dense) or N*K (sparse), doubled for complex
    const double cnt = (double)g.N * (g.sparse ? g.K : 1) * (CPLX ? 2 : 1);
    int log2cnt = 0;
    while ((double)(1ll << log2cnt) < cnt && log2cnt < 62) log2cnt++;

    {
    ProfScope pt(st, "transpose_gout");
    dev::fill_zero_kernel<<<1, 64, 0, st>>>(reinterpret_cast<float *>(maxes), 2);
    dev::transpose_gout_kernel<ACC64, CPLX, FS><<<dim3((g.N + 255) / 256, nchunks), 256, 0, st>>>(
        gout, gs_b, g.B, g.N, ax, static_cast<A *>(gtr), g, perm, maxes);
    }
    dim3 grid(g.C * g.T, nchunks);
    const int64_t gx_t = g.S, gx_c = (int64_t)g.T * g.S, gx_b = (int64_t)g.C * g.T * g.S;
    (void)hipFuncSetAttribute((const void *)dev::das_bwd_x_kernel<ACC64, CPLX, PRECISE, FS>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    ProfScope ps(st, "das_bwd_x");
    dev::Geo ga = g;
    ga.dstpos = perm;
    ga.offsets = perm + 3 * (int64_t)g.N;
    ga.weights = ga.offsets + (int64_t)g.K * g.N;
    ga.sparse = g.sparse ? reinterpret_cast<const int *>(ga.weights + (int64_t)g.K * g.N) : nullptr;
    // sparse: inverted index after the geometry, [count T][start T+1][list K*N][cursor T or
    // block counts T*nblk]
    unsigned *sp_start = nullptr;
    int *sp_list = nullptr;
    if (g.sparse && g.T <= dev::kMaxSparseT) {
        unsigned *tot = reinterpret_cast<unsigned *>(perm + (3 + 3 * (int64_t)g.K) * g.N);
        sp_start = tot + g.T;
        sp_list = reinterpret_cast<int *>(sp_start + g.T + 1);
        unsigned *tc = reinterpret_cast<unsigned *>(sp_list + (int64_t)g.K * g.N);
        const int nblk = (g.N + dev::kSpBlock - 1) / dev::kSpBlock;
        dev::sparse_bcount_kernel<<<nblk, 256, 0, st>>>(ga.sparse, g.K, g.N, g.T, nblk, tc);
        dev::sparse_bscan_kernel<<<g.T, 256, 0, st>>>(tc, nblk, tot);
        dev::sparse_scan_kernel<<<1, 64, 0, st>>>(tot, g.T, sp_start);
        dev::sparse_bfill_kernel<<<nblk, 256, 0, st>>>(ga.sparse, g.K, g.N, g.T, nblk, tc,
                                                        sp_start, sp_list);
    } else if (g.sparse) {
        unsigned *count = reinterpret_cast<unsigned *>(perm + (3 + 3 * (int64_t)g.K) * g.N);
        sp_start = count + g.T;
        sp_list = reinterpret_cast<int *>(sp_start + g.T + 1);
        unsigned *cursor = reinterpret_cast<unsigned *>(sp_list + (int64_t)g.K * g.N);
        const int64_t entries = (int64_t)g.K * g.N;
        dev::fill_zero_kernel<<<(g.T + 255) / 256, 256, 0, st>>>(reinterpret_cast<float *>(count), g.T);
        dev::fill_zero_kernel<<<(g.T + 255) / 256, 256, 0, st>>>(reinterpret_cast<float *>(cursor), g.T);
        const int cgrid = (int)std::min<int64_t>((entries + 255) / 256, 2048);
        dev::sparse_count_kernel<<<cgrid, 256, 0, st>>>(g.sparse, entries, g.T, count);
        dev::sparse_scan_kernel<<<1, 64, 0, st>>>(count, g.T, sp_start);
        if ((entries + 255) / 256 > 0x7fffffff) return hipErrorInvalidValue;
        dev::sparse_fill_kernel<<<(unsigned)((entries + 255) / 256), 256, 0, st>>>(
            g.sparse, g.K, g.N, g.T, ax, sp_start, cursor, sp_list);
    }
    if (const char *env = std::getenv("FFDAS_BWDX_THREADS")) threads = std::atoi(env);
    if (threads != 256 && threads != 512 && threads != 1024) threads = 1024;
    const int rows = g.C * g.T;
    int tail = 0;
    if constexpr (FS >= 4) {
        const char *env = std::getenv("FFDAS_BWDX_TAIL");
        if (!g.sparse && !(env && env[0] == '0')) {
            const auto kh = dev::das_bwd_x_kernel<ACC64, CPLX, PRECISE, FS / 2>;
            const size_t lds_h = bwdx_lds_bytes(g.S, FS / 2, CPLX, ACC64);
            (void)hipFuncSetAttribute((const void *)kh, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds_h);
            int per_cu = 0, per_cu_h = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                    &per_cu, (const void *)dev::das_bwd_x_kernel<ACC64, CPLX, PRECISE, FS>,
                    threads, lds) == hipSuccess &&
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_h, (const void *)kh, threads,
                                                             lds_h) == hipSuccess)
                tail = bwdx_tail_rows(rows, nchunks, cu_count, per_cu, per_cu_h);
            if (tail > 0)
                kh<<<dim3(tail, (g.B + FS / 2 - 1) / (FS / 2)), threads, lds_h, st>>>(
                    ga, ax, gtr, maxes, log2cnt, gx, gx_b, gx_c, gx_t, sp_start, sp_list,
                    rows - tail, FS);
        }
    }
    grid.x = rows - tail;
    dev::das_bwd_x_kernel<ACC64, CPLX, PRECISE, FS><<<grid, threads, lds, st>>>(
        ga, ax, gtr, maxes, log2cnt, gx, gx_b, gx_c, gx_t, sp_start, sp_list, 0, FS);
    return hipGetLastError();
}

template <bool ACC64, bool CPLX, bool PRECISE>
hipError_t launch_bwdx_t(hipStream_t st, const dev::Geo &g, const dev::Axial &tl, int threads,
                         const void *gout, int64_t gs_b, void *gt, void *gx, int FS, int cus) {
    switch (FS) {
        case 16: return launch_bwdx_fs<ACC64, CPLX, PRECISE, 16>(st, g, tl, threads, gout, gs_b, gt, gx, cus);
        case 8: return launch_bwdx_fs<ACC64, CPLX, PRECISE, 8>(st, g, tl, threads, gout, gs_b, gt, gx, cus);
        case 4: return launch_bwdx_fs<ACC64, CPLX, PRECISE, 4>(st, g, tl, threads, gout, gs_b, gt, gx, cus);
        case 2: return launch_bwdx_fs<ACC64, CPLX, PRECISE, 2>(st, g, tl, threads, gout, gs_b, gt, gx, cus);
        default: return launch_bwdx_fs<ACC64, CPLX, PRECISE, 1>(st, g, tl, threads, gout, gs_b, gt, gx, cus);
    }
}

// Frame-inner layout of a problem: frame-group planes for dense problems, whole rows for
// sparse compounding (see FrameInnerRows).  On the 2-D benchmark (dense) the planes took the
// FP16 forward from 0.268 to 0.219 ms; on the batched sparse benchmark the rows fetch 2.56 TB
// per ensemble beyond L2 against 3.44 TB for the planes, 498 against 574 ms.
bool fi_planes(const DasProblem &p) { return !p.sparse; }
bool fi_planes(const dev::Geo &g) { return g.sparse == nullptr; }

// Taps whose gathers a frame-inner forward wave issues together (1, 2 or 4).  Measured on
// the 2-D fwd+adjoint benchmark (16 frames per chunk, frame-group planes): FP16 storage
// 0.280 ms (1 tap), 0.230 (2), 0.219 (4); FP32 storage (row-pair layout) 0.494 / 0.436 /
// 0.429 ms.  Complex data at 4 taps spills registers.  The parameter adjoint keeps one tap
// per group: 2 and 4 measured 0.302 and 0.292 ms against 0.238 (its registers drop it to 3
// and 2 waves per SIMD).  FFDAS_TAPS overrides.
template <Storage ST, bool CPLX>
int tap_group() {
    static const int env = [] {
        const char *e = std::getenv("FFDAS_TAPS");
        return e ? std::atoi(e) : 0;
    }();
    if (env > 0) return env >= 4 && !CPLX ? 4 : (env >= 2 ? 2 : 1);
    return CPLX ? 2 : 4;
}

template <Storage ST, bool CPLX, int F>
constexpr bool fi_ok() {
    return (F * (CPLX ? 2 : 1) * (ST == Storage::F32 ? 4 : 2)) % 16 == 0;
}

template <Storage ST, bool CPLX, int F>
hipError_t launch_fwd_fi_f(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl, int groups,
                           const void *x, void *out, int64_t os_b, double br, double bi) {
    if constexpr (!fi_ok<ST, CPLX, F>()) {
        return hipErrorInvalidValue;
    } else {
        constexpr bool PRECISE = true;
        const int ntiles = tl.nta * tl.nth * tl.ntw;
        dim3 grid(ntiles, (g.B + F - 1) / F);
        const int G = channel_groups(g, (int64_t)grid.x * grid.y, groups);
        const size_t lds = G > 1 ? (size_t)G * F * (CPLX ? 2 : 1) * 64 * sizeof(float) : 0;
        auto go = [&](auto kern) {
            if (lds > 0)
                (void)hipFuncSetAttribute((const void *)kern,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            ProfScope ps(st, "das_fwd_fi");
            kern<<<grid, 64 * G, lds, st>>>(g, tl, x, out, os_b, br, bi);
            return hipGetLastError();
        };
        auto by_taps = [&](auto pl) {
            constexpr bool PL = decltype(pl)::value;
            const int tb = tap_group<ST, CPLX>();
            if constexpr (!CPLX)
                if (tb == 4) return go(dev::das_fwd_fi_kernel<ST, false, CPLX, PRECISE, F, kKC, 4, PL>);
            if (tb >= 2) return go(dev::das_fwd_fi_kernel<ST, false, CPLX, PRECISE, F, kKC, 2, PL>);
            return go(dev::das_fwd_fi_kernel<ST, false, CPLX, PRECISE, F, kKC, 1, PL>);
        };
        return fi_planes(g) ? by_taps(std::true_type{}) : by_taps(std::false_type{});
    }
}

template <Storage ST, bool CPLX>
hipError_t launch_fwd_fi_t(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl, int F,
                           int groups, const void *x, void *out, int64_t os_b, double br,
                           double bi) {
    switch (F) {
        case 16: return launch_fwd_fi_f<ST, CPLX, 16>(st, g, tl, groups, x, out, os_b, br, bi);
        case 8: return launch_fwd_fi_f<ST, CPLX, 8>(st, g, tl, groups, x, out, os_b, br, bi);
        case 4: return launch_fwd_fi_f<ST, CPLX, 4>(st, g, tl, groups, x, out, os_b, br, bi);
        default: return hipErrorInvalidValue;
    }
}

template <Storage ST, bool CPLX, int F>
hipError_t launch_bwdp_fi_f(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl, int groups,
                            const void *x, const void *gout, int64_t gs_b, const ParamOut &po) {
    if constexpr (!fi_ok<ST, CPLX, F>()) {
        return hipErrorInvalidValue;
    } else {
        constexpr bool PRECISE = true;
        const int ntiles = tl.nta * tl.nth * tl.ntw;
        dim3 grid(ntiles, (g.B + F - 1) / F);
        const int G = channel_groups(g, (int64_t)grid.x * grid.y, groups);
        const size_t lds = G > 1 ? (size_t)G * 2 * kKC * 64 * sizeof(float) : 0;
        auto go = [&](auto kern) {
            if (lds > 0)
                (void)hipFuncSetAttribute((const void *)kern,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            ProfScope ps(st, "das_bwd_param_fi");
            kern<<<grid, 64 * G, lds, st>>>(
                g, tl, x, gout, gs_b, grid.y > 1 ? po.pofs : po.gofs, grid.y > 1 ? po.pwt : po.gwt,
                grid.y > 1 ? (int64_t)g.K * g.N : 0, po.gscale ? po.spart : nullptr);
            return hipGetLastError();
        };
        const hipError_t e = fi_planes(g)
                                 ? go(dev::das_bwd_param_fi_kernel<ST, false, CPLX, PRECISE, F, kKC, true>)
                                 : go(dev::das_bwd_param_fi_kernel<ST, false, CPLX, PRECISE, F, kKC, false>);
        if (e != hipSuccess) return e;
        return finish_params(st, po, (int)grid.y, (int64_t)g.K * g.N, (int)(grid.x * grid.y));
    }
}

template <Storage ST, bool CPLX>
hipError_t launch_bwdp_fi_t(hipStream_t st, const dev::Geo &g, const dev::Tiling &tl, int F,
                            int groups, const void *x, const void *gout, int64_t gs_b,
                            const ParamOut &po) {
    switch (F) {
        case 16: return launch_bwdp_fi_f<ST, CPLX, 16>(st, g, tl, groups, x, gout, gs_b, po);
        case 8: return launch_bwdp_fi_f<ST, CPLX, 8>(st, g, tl, groups, x, gout, gs_b, po);
        case 4: return launch_bwdp_fi_f<ST, CPLX, 4>(st, g, tl, groups, x, gout, gs_b, po);
        default: return hipErrorInvalidValue;
    }
}

template <Storage ST, bool CPLX, int F>
hipError_t launch_fwd_batched_f(hipStream_t st, const std::vector<BatchMember> &ms) {
    if constexpr (!fi_ok<ST, CPLX, F>()) {
        return hipErrorInvalidValue;
    } else {
        constexpr bool PRECISE = true;
        for (size_t off = 0; off < ms.size(); off += dev::kMaxBatch) {
            dev::BatchArgs args{};
            const int n = (int)std::min(ms.size() - off, (size_t)dev::kMaxBatch);
            int64_t start = 0;
            int minC = 1 << 30;
            int want = 0;
            for (int i = 0; i < n; i++) {
                const BatchMember &m = ms[off + i];
                dev::BatchItem &it = args.item[i];
                it.g = make_geo(m.a, m.p);
                it.tl = make_tiling(m.p.spatial, m.p.tune);
                const int tiles = it.tl.nta * it.tl.nth * it.tl.ntw;
                const int chunks = (m.p.B + F - 1) / F;
                it.x = m.xfi;
                it.out = m.out;
                it.os_b = m.p.os_b;
                it.beta_re = m.beta_re;
                it.beta_im = m.beta_im;
                it.start = (int)start;
                it.tiles = tiles;
                start += (int64_t)tiles * chunks;
                minC = std::min(minC, m.p.C);
                if (i == 0) want = m.p.tune.groups;
            }
            args.count = n;
            args.xcd = xcd_groups() ? 1 : 0;
            if (start == 0) continue;
            if (start > 0x7fffffff) return hipErrorInvalidValue;
            dev::Geo gc{};
            gc.C = minC;
            const int G = channel_groups(gc, start, want);
            const size_t lds = G > 1 ? (size_t)G * F * (CPLX ? 2 : 1) * 64 * sizeof(float) : 0;
            auto go = [&](auto kern) {
                if (lds > 0)
                    (void)hipFuncSetAttribute((const void *)kern,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                ProfScope ps(st, "das_fwd_fi_batched");
                kern<<<(unsigned)start, 64 * G, lds, st>>>(args);
                return hipGetLastError();
            };
            // the members of one launch share the layout (the caller groups them so)
            auto by_taps = [&](auto pl) {
                constexpr bool PL = decltype(pl)::value;
                const int tb = tap_group<ST, CPLX>();
                if constexpr (!CPLX)
                    if (tb == 4) return go(dev::das_fwd_fi_batched_kernel<ST, CPLX, PRECISE, F, kKC, 4, PL>);
                if (tb >= 2) return go(dev::das_fwd_fi_batched_kernel<ST, CPLX, PRECISE, F, kKC, 2, PL>);
                return go(dev::das_fwd_fi_batched_kernel<ST, CPLX, PRECISE, F, kKC, 1, PL>);
            };
            const hipError_t e = fi_planes(ms[off].p) ? by_taps(std::true_type{})
                                                      : by_taps(std::false_type{});
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
}

template <Storage ST, bool CPLX>
hipError_t launch_fwd_batched_t(hipStream_t st, const std::vector<BatchMember> &ms, int F) {
    switch (F) {
        case 16: return launch_fwd_batched_f<ST, CPLX, 16>(st, ms);
        case 8: return launch_fwd_batched_f<ST, CPLX, 8>(st, ms);
        case 4: return launch_fwd_batched_f<ST, CPLX, 4>(st, ms);
        default: return hipErrorInvalidValue;
    }
}

template <Storage SS, class DST, bool CPLX, int F>
hipError_t launch_relayout_f(hipStream_t st, const DasProblem &p, const void *x, void *dst) {
    constexpr int CW = CPLX ? 2 : 1;
    constexpr size_t ES = sizeof(typename dev::Elem<SS>::raw) * CW;   // bytes per sample
    const bool vec = (reinterpret_cast<uintptr_t>(x) % 16 == 0) && (p.xs_b * ES) % 16 == 0 &&
                     (p.xs_c * ES) % 16 == 0 && (p.xs_t * ES) % 16 == 0 &&
                     ((size_t)p.S * ES) % 16 == 0 && p.S % 2 == 0 &&
                     (F * CW * sizeof(DST)) % 16 == 0;
    // frame groups of the layout the gathers read: planes for dense problems, whole rows
    // for sparse compounding (FrameInnerRows)
    auto launch = [&](auto gs) {
        constexpr int GS = decltype(gs)::value;
        if (vec) {
            const int64_t blocks = (int64_t)((p.S + dev::kRelL - 1) / dev::kRelL) * p.C * p.T;
            if (blocks > 0x7fffffff) return hipErrorInvalidValue;
            dim3 grid((unsigned)blocks, (p.B + F - 1) / F);
            ProfScope ps(st, "relayout_fi");
            dev::relayout_fi_vec_kernel<SS, DST, F, CW, GS><<<grid, 256, 0, st>>>(
                static_cast<const typename dev::Elem<SS>::raw *>(x), p.xs_b, p.xs_c, p.xs_t, p.B,
                p.C, p.T, p.S, static_cast<DST *>(dst));
            return hipGetLastError();
        }
        const int64_t blocks = (int64_t)((p.S + 63) / 64) * p.C * p.T;
        if (blocks > 0x7fffffff) return hipErrorInvalidValue;
        dim3 grid((unsigned)blocks, (p.B + F - 1) / F);
        ProfScope ps(st, "relayout_fi");
        dev::relayout_fi_kernel<SS, DST, F, CW, GS><<<grid, 256, 0, st>>>(
            static_cast<const typename dev::Elem<SS>::raw *>(x), p.xs_b, p.xs_c, p.xs_t, p.B, p.C,
            p.T, p.S, static_cast<DST *>(dst));
        return hipGetLastError();
    };
    constexpr int ESZ = (int)sizeof(DST);
    if (fi_planes(p))
        return launch(std::integral_constant<int, dev::fi_group<F, CW, ESZ, true>()>{});
    return launch(std::integral_constant<int, dev::fi_group<F, CW, ESZ, false>()>{});
}

template <Storage SS, class DST, bool CPLX>
hipError_t launch_relayout_t(hipStream_t st, const DasProblem &p, const void *x, void *dst) {
    switch (p.fi_frames) {
        case 16: return launch_relayout_f<SS, DST, CPLX, 16>(st, p, x, dst);
        case 8: return launch_relayout_f<SS, DST, CPLX, 8>(st, p, x, dst);
        case 4: return launch_relayout_f<SS, DST, CPLX, 4>(st, p, x, dst);
        default: return hipErrorInvalidValue;
    }
}

// ---- grad_x by bucketing ----
// Frames per workgroup for the bucketed grad_x kernel, limited by the accumulator registers
// ((KPT+1) * F * CW floats per thread).
int sorted_frames(int KPT, bool cplx, int B) {
    // largest frame count whose accumulators stay in registers (no scratch) at 1024 threads
    int fmax = cplx ? (KPT >= 4 ? 4 : 8) : (KPT >= 2 ? 8 : 16);
    int want = B >= 16 ? 16 : B >= 8 ? 8 : 4;
    return want < fmax ? want : fmax;
}

size_t sorted_lds_bytes(int S, int F, bool cplx) {
    const int CW = cplx ? 2 : 1;
    const size_t rec = cplx ? 20 : 12;
    return (size_t)(S + 1) * 4 + (size_t)(dev::kSortNT / 64) * 4 +
           (size_t)(dev::kSortNT / 64) * F * CW * 4 +
           (size_t)(cplx ? dev::sort_vpt<true>() : dev::sort_vpt<false>()) * dev::kSortNT * rec;
}

template <bool CPLX, bool PRECISE, int KPT, int F>
hipError_t launch_bwdx_sorted_kf(hipStream_t st, const dev::Geo &g, const void *gout, int64_t gs_b,
                                 void *ws, void *gx) {
    if constexpr ((CPLX && F * (KPT + 1) > 24) || (!CPLX && F * (KPT + 1) > 48)) {
        return hipErrorInvalidValue;
    } else {
        constexpr int CW = CPLX ? 2 : 1;
        const int nchunks = (g.B + F - 1) / F;
        float *gt = static_cast<float *>(ws);
        {
            ProfScope pt(st, "transpose_gout");
            dev::transpose_gout_f_kernel<CPLX, F><<<dim3((g.N + 255) / 256, nchunks), 256, 0, st>>>(
                static_cast<const float *>(gout), gs_b, g.B, g.N, gt);
        }
        const int kslots = g.sparse ? g.K : 1;
        const int nc = std::max(1, dev::sort_vpt<CPLX>() * dev::kSortNT / kslots);
        const size_t lds = sorted_lds_bytes(g.S, F, CPLX);
        dim3 grid(g.C * g.T, nchunks);
        const int64_t gx_t = g.S, gx_c = (int64_t)g.T * g.S, gx_b = (int64_t)g.C * g.T * g.S;
        (void)hipFuncSetAttribute((const void *)dev::das_bwd_x_sorted_kernel<CPLX, PRECISE, KPT, F>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        ProfScope ps(st, "das_bwd_x");
        dev::das_bwd_x_sorted_kernel<CPLX, PRECISE, KPT, F><<<grid, dev::kSortNT, lds, st>>>(
            g, gt, static_cast<float *>(gx), gx_b, gx_c, gx_t, nc);
        (void)CW;
        return hipGetLastError();
    }
}

template <bool CPLX, bool PRECISE, int KPT>
hipError_t launch_bwdx_sorted_k(hipStream_t st, const dev::Geo &g, const void *gout, int64_t gs_b,
                                void *ws, void *gx, int F) {
    switch (F) {
        case 16: return launch_bwdx_sorted_kf<CPLX, PRECISE, KPT, 16>(st, g, gout, gs_b, ws, gx);
        case 8: return launch_bwdx_sorted_kf<CPLX, PRECISE, KPT, 8>(st, g, gout, gs_b, ws, gx);
        default: return launch_bwdx_sorted_kf<CPLX, PRECISE, KPT, 4>(st, g, gout, gs_b, ws, gx);
    }
}

template <bool CPLX, bool PRECISE>
hipError_t launch_bwdx_sorted_t(hipStream_t st, const dev::Geo &g, const void *gout, int64_t gs_b,
                                void *ws, void *gx, int KPT, int F) {
    switch (KPT) {
        case 1: return launch_bwdx_sorted_k<CPLX, PRECISE, 1>(st, g, gout, gs_b, ws, gx, F);
        case 2: return launch_bwdx_sorted_k<CPLX, PRECISE, 2>(st, g, gout, gs_b, ws, gx, F);
        default: return launch_bwdx_sorted_k<CPLX, PRECISE, 4>(st, g, gout, gs_b, ws, gx, F);
    }
}

}  // namespace

ffdas_error_t launch_forward_batched(ffdas_context &h, const std::vector<BatchMember> &ms) {
    if (ms.empty()) return FFDAS_SUCCESS;
    const DasProblem &p = ms[0].p;
    hipError_t e = hipErrorInvalidValue;
    switch (p.k_storage) {
        case Storage::F16:
            e = p.cplx ? launch_fwd_batched_t<Storage::F16, true>(h.stream, ms, p.fi_frames)
                       : launch_fwd_batched_t<Storage::F16, false>(h.stream, ms, p.fi_frames);
            break;
        case Storage::BF16:
            e = p.cplx ? launch_fwd_batched_t<Storage::BF16, true>(h.stream, ms, p.fi_frames)
                       : launch_fwd_batched_t<Storage::BF16, false>(h.stream, ms, p.fi_frames);
            break;
        case Storage::F32:
            e = p.cplx ? launch_fwd_batched_t<Storage::F32, true>(h.stream, ms, p.fi_frames)
                       : launch_fwd_batched_t<Storage::F32, false>(h.stream, ms, p.fi_frames);
            break;
        default: return FFDAS_ERROR_INTERNAL;
    }
    return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

ffdas_error_t launch_forward(ffdas_context &h, const DasProblem &p, const DasGeomArgs &a,
                             const XView &xv, void *out, double beta_re, double beta_im) {
    const dev::Geo g = make_geo(a, p);
    const dev::Tiling tl = make_tiling(p.spatial, p.tune);
    hipError_t e = hipSuccess;
    hipStream_t st = h.stream;
    const int64_t os_b = p.os_b;
    if (p.fi_frames > 0) {
        const int F = p.fi_frames;
        switch (p.k_storage) {
            case Storage::F16:
                e = p.cplx ? launch_fwd_fi_t<Storage::F16, true>(st, g, tl, F, p.tune.groups, xv.ptr, out, os_b, beta_re, beta_im)
                           : launch_fwd_fi_t<Storage::F16, false>(st, g, tl, F, p.tune.groups, xv.ptr, out, os_b, beta_re, beta_im);
                break;
            case Storage::BF16:
                e = p.cplx ? launch_fwd_fi_t<Storage::BF16, true>(st, g, tl, F, p.tune.groups, xv.ptr, out, os_b, beta_re, beta_im)
                           : launch_fwd_fi_t<Storage::BF16, false>(st, g, tl, F, p.tune.groups, xv.ptr, out, os_b, beta_re, beta_im);
                break;
            case Storage::F32:
                e = p.cplx ? launch_fwd_fi_t<Storage::F32, true>(st, g, tl, F, p.tune.groups, xv.ptr, out, os_b, beta_re, beta_im)
                           : launch_fwd_fi_t<Storage::F32, false>(st, g, tl, F, p.tune.groups, xv.ptr, out, os_b, beta_re, beta_im);
                break;
            default: return FFDAS_ERROR_INTERNAL;
        }
        return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
    }
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

// Frames per workgroup of the grad_x scatter: the largest of 16/8/4/2/1 whose LDS row fits
// in a CU's LDS.  Measured on gfx950, one workgroup of 16 frames per CU beats two of 8: the
// per-voxel delay arithmetic and loads are shared by more frames.  FFDAS_BWDX_FRAMES
// overrides.
int bwdx_frames(const DasProblem &p) {
    auto lds = [&](int fs) { return bwdx_lds_bytes(p.S, fs, p.cplx, p.acc64); };
    if (const char *env = std::getenv("FFDAS_BWDX_FRAMES")) {
        const int fs = std::atoi(env);
        if ((fs == 1 || fs == 2 || fs == 4 || fs == 8 || fs == 16) && lds(fs) <= (size_t)kLdsLimit)
            return fs;
    }
    int cap = p.B >= 16 ? 16 : p.B >= 8 ? 8 : p.B >= 4 ? 4 : p.B >= 2 ? 2 : 1;
    if (p.cplx && cap > 8) cap = 8;   // 16 complex frames exceed the register budget
    const int want = p.tune.bwdx_frames;
    if ((want == 1 || want == 2 || want == 4 || want == 8 || want == 16) && want <= cap &&
        lds(want) <= (size_t)kLdsLimit)
        return want;
    for (int fs = cap; fs >= 1; fs /= 2)
        if (lds(fs) <= (size_t)kLdsLimit) return fs;
    return 0;
}

// The bucketed grad_x kernel (FFDAS_BWDX=sorted; float data, S <= 4096) is selectable; the
// default is the fixed-point LDS-atomic kernel, which is faster on the measured workloads and
// bitwise reproducible.
bool use_sorted_bwdx(const DasProblem &p) {
    if (p.acc64 || p.S > 4096) return false;
    const char *env = std::getenv("FFDAS_BWDX");
    return env && std::strcmp(env, "sorted") == 0;
}

int sorted_kpt(const DasProblem &p) { return p.S <= 1024 ? 1 : (p.S <= 2048 ? 2 : 4); }

size_t backward_workspace(const DasProblem &p) {
    if (p.plain) return 0;
    if (use_sorted_bwdx(p)) {
        const int F = sorted_frames(sorted_kpt(p), p.cplx, p.B);
        const size_t chunks = (size_t)(p.B + F - 1) / F;
        return chunks * (size_t)F * (size_t)p.N * (p.cplx ? 2 : 1) * sizeof(float);
    }
    const int fs = bwdx_frames(p);
    if (fs <= 0) return 0;
    const size_t esz = 8 * (p.cplx ? 2 : 1);   // pre-scaled double grad_out
    const size_t nchunks = (size_t)(p.B + fs - 1) / fs;
    const size_t perm = (size_t)p.N * (3 + 3 * (size_t)p.K) * sizeof(float);   // axial geometry
    // [count T][start T+1][list K*N][cursor T or block counts T * ceil(N / 1024)]
    const size_t nblk = ((size_t)p.N + 1023) / 1024;
    const size_t inv =
        p.sparse ? ((2 * (size_t)p.T + 1) + (size_t)p.K * p.N + (size_t)p.T * std::max<size_t>(nblk, 1)) * 4
                 : 0;
    return 256 + nchunks * (size_t)fs * (size_t)p.N * esz + perm + inv;
}

// Partial buffers of the parameter adjoint (see ParamOut), carved from `pws`; the layout
// matches param_workspace().
namespace {
struct ParamLayout {
    int chunks = 1;
    int blocks = 0;
    size_t ofs_bytes = 0, wt_bytes = 0, scale_bytes = 0;
};
size_t align256(size_t v) { return (v + 255) & ~size_t(255); }
ParamLayout param_layout(const DasProblem &p, bool has_o, bool has_w, bool has_s) {
    ParamLayout L;
    const dev::Tiling tl = make_tiling(p.spatial, p.tune);
    const int ntiles = tl.nta * tl.nth * tl.ntw;
    const int FT = p.fi_frames > 0 ? p.fi_frames : frame_tile_for(p.B);
    L.chunks = (p.B + FT - 1) / FT;
    L.blocks = (p.fi_frames > 0 ? ntiles : (ntiles + 3) / 4) * L.chunks;
    const size_t slice = (size_t)L.chunks * (size_t)p.K * (size_t)p.N * sizeof(float);
    if (L.chunks > 1) {
        L.ofs_bytes = has_o ? align256(slice) : 0;
        L.wt_bytes = has_w ? align256(slice) : 0;
    }
    L.scale_bytes = has_s ? align256((size_t)L.blocks * sizeof(double)) : 0;
    return L;
}
ParamOut param_out(const DasProblem &p, float *gofs, float *gwt, float *gscale, void *pws) {
    const ParamLayout L = param_layout(p, gofs != nullptr, gwt != nullptr, gscale != nullptr);
    ParamOut po;
    po.gofs = gofs;
    po.gwt = gwt;
    po.gscale = gscale;
    char *b = static_cast<char *>(pws);
    if (L.ofs_bytes) po.pofs = reinterpret_cast<float *>(b);
    b += L.ofs_bytes;
    if (L.wt_bytes) po.pwt = reinterpret_cast<float *>(b);
    b += L.wt_bytes;
    if (L.scale_bytes) po.spart = reinterpret_cast<double *>(b);
    return po;
}
}  // namespace

size_t param_workspace(const DasProblem &p, bool has_o, bool has_w, bool has_s) {
    if (!has_o && !has_w && !has_s) return 0;
    const ParamLayout L = param_layout(p, has_o, has_w, has_s);
    return L.ofs_bytes + L.wt_bytes + L.scale_bytes;
}

ffdas_error_t launch_backward(ffdas_context &h, const DasProblem &p, const DasGeomArgs &a,
                              const XView &xv, const void *gout, int64_t gs_b, void *gx,
                              float *gofs, float *gwt, float *gscale, void *gt, void *pws) {
    const dev::Geo g = make_geo(a, p);
    const dev::Tiling tl = make_tiling(p.spatial, p.tune);
    hipStream_t st = h.stream;
    hipError_t e = hipSuccess;

    // When both the parameter gradients and grad_x are asked for with 16-bit storage, the
    // grad_x scatter (bound by LDS atomics) runs on the handle's second stream beside the
    // parameter adjoint (bound by its gathers), forked from and joined back into the
    // handle's stream: a grad_x workgroup leaves registers and waves for a parameter-adjoint
    // wave on the same CU, and the two use different units.  Measured on the 2-D benchmark:
    // FP16 storage 0.994-0.998 -> 0.976-0.979 ms a step; FP32 storage 1.407 -> 1.437 ms (its
    // heavier gathers lose more to the sharing than the overlap gains), so 32- and 64-bit
    // storage stay serial.  FFDAS_BWD_CONCURRENT=0 / 1 forces serial / concurrent.
    const bool params = gofs || gwt || gscale;
    const bool can = params && gx && !p.plain && !use_sorted_bwdx(p);
    bool concurrent = can && (p.k_storage == Storage::F16 || p.k_storage == Storage::BF16);
    if (const char *env = std::getenv("FFDAS_BWD_CONCURRENT"))
        concurrent = can && std::atoi(env) != 0;
    hipStream_t sx = st;
    if (concurrent) {
        FFDAS_TRY(h.side_stream(&sx));
        if ((e = hipEventRecord(h.fork_ev, st)) != hipSuccess ||
            (e = hipStreamWaitEvent(sx, h.fork_ev, 0)) != hipSuccess)
            return hip_status(e);
    }

    if (params) {
        const ParamOut po = param_out(p, gofs, gwt, gscale, pws);
#define BWDP(ST, A64, CX) e = launch_bwdp_t<ST, A64, CX>(st, g, tl, xv, gout, gs_b, po)
#define BWDPFI(ST, CX) e = launch_bwdp_fi_t<ST, CX>(st, g, tl, p.fi_frames, p.tune.param_groups, xv.ptr, gout, gs_b, po)
        if (p.fi_frames > 0) {
            switch (p.k_storage) {
                case Storage::F16: if (p.cplx) BWDPFI(Storage::F16, true); else BWDPFI(Storage::F16, false); break;
                case Storage::BF16: if (p.cplx) BWDPFI(Storage::BF16, true); else BWDPFI(Storage::BF16, false); break;
                case Storage::F32: if (p.cplx) BWDPFI(Storage::F32, true); else BWDPFI(Storage::F32, false); break;
                default: return FFDAS_ERROR_INTERNAL;
            }
        } else if (!p.cplx) {
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
#undef BWDPFI
        if (e != hipSuccess) return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
    }
    if (gx && p.plain) {
        const int64_t cnt = (int64_t)p.B * p.C * p.T * p.S * (p.cplx ? 2 : 1) * (p.acc64 ? 2 : 1);
        FFDAS_TRY(launch_fill_zero(h, static_cast<float *>(gx), cnt));
        if (p.acc64) {
            e = p.cplx ? launch_bwdx_plain_t<true, true, true>(st, g, tl, gout, gs_b, gx)
                       : launch_bwdx_plain_t<true, false, true>(st, g, tl, gout, gs_b, gx);
        } else {
            e = p.cplx ? launch_bwdx_plain_t<false, true, true>(st, g, tl, gout, gs_b, gx)
                       : launch_bwdx_plain_t<false, false, true>(st, g, tl, gout, gs_b, gx);
        }
        if (e != hipSuccess) return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
    } else if (gx && use_sorted_bwdx(p)) {
        const int KPT = sorted_kpt(p);
        const int F = sorted_frames(KPT, p.cplx, p.B);
        e = p.cplx ? launch_bwdx_sorted_t<true, true>(st, g, gout, gs_b, gt, gx, KPT, F)
                   : launch_bwdx_sorted_t<false, true>(st, g, gout, gs_b, gt, gx, KPT, F);
        if (e != hipSuccess) return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
    } else if (gx) {
        const int FS = bwdx_frames(p);
        if (FS <= 0) return FFDAS_ERROR_DIMS_TOO_LARGE;
        const dev::Axial ax = make_axial(p);
        // grad_x never reads the channel data, so its delays are the compensated FP32 ones
        // whatever the storage type: the scatter is bound by LDS atomics, and the extra
        // delay arithmetic overlaps with them
        if (p.acc64) {
            e = p.cplx ? launch_bwdx_t<true, true, true>(sx, g, ax, p.tune.bwdx_threads, gout, gs_b, gt, gx, FS, h.cu_count)
                       : launch_bwdx_t<true, false, true>(sx, g, ax, p.tune.bwdx_threads, gout, gs_b, gt, gx, FS, h.cu_count);
        } else {
            e = p.cplx ? launch_bwdx_t<false, true, true>(sx, g, ax, p.tune.bwdx_threads, gout, gs_b, gt, gx, FS, h.cu_count)
                       : launch_bwdx_t<false, false, true>(sx, g, ax, p.tune.bwdx_threads, gout, gs_b, gt, gx, FS, h.cu_count);
        }
        if (e != hipSuccess) return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
    }
    if (sx != st) {
        if ((e = hipEventRecord(h.join_ev, sx)) != hipSuccess ||
            (e = hipStreamWaitEvent(st, h.join_ev, 0)) != hipSuccess)
            return hip_status(e);
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
    ProfScope ps(st, "cast_rows");
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

// Frame-inner layout choice: frames per chunk F (16/8/4) for the gather kernels, or 0 for the
// natural layout.  ALG1 keeps the natural-layout (plain) kernels; double accumulation and
// double storage keep them too.  Padding may add at most as many frames as there are.
int fi_frames_for(const DasProblem &p, ffdas_alg_t alg) {
    if (alg == FFDAS_ALG1 || p.acc64 || p.k_storage == Storage::F64 || p.B < 4) return 0;
    if (const char *env = std::getenv("FFDAS_LAYOUT"))
        if (std::strcmp(env, "natural") == 0) return 0;
    const int esz = (int)storage_bytes(p.k_storage) * (p.cplx ? 2 : 1);
    // the gathers address one channel's rows ((T, S, F) samples) with 32-bit offsets
    auto fits = [&](int F) { return (double)p.T * p.S * F * esz < 2147483648.0 - 64.0; };
    const int want = p.tune.fi_frames;
    if ((want == 4 || want == 8 || want == 16) && (want * esz) % 16 == 0 && fits(want) &&
        (double)((p.B + want - 1) / want) * want <= 2.0 * p.B)
        return want;
    // cost ~ padded frames * (1 + G / F): per-voxel delay arithmetic is worth about G
    // interpolations and is shared by the F frames of a chunk
    constexpr double G = 10.0;
    int best = 0;
    double best_cost = 0.0;
    for (int F : {16, 8, 4}) {
        if ((F * esz) % 16 != 0 || !fits(F)) continue;
        const double padded = (double)((p.B + F - 1) / F) * F;
        if (padded > 2.0 * p.B) continue;
        const double cost = padded * (1.0 + G / F);
        if (best == 0 || cost < best_cost) {
            best = F;
            best_cost = cost;
        }
    }
    return best;
}

size_t fi_bytes(const DasProblem &p) {
    if (p.fi_frames <= 0) return 0;
    const size_t chunks = (size_t)(p.B + p.fi_frames - 1) / p.fi_frames;
    return chunks * (size_t)p.fi_frames * (size_t)p.C * p.T * p.S * (p.cplx ? 2 : 1) *
           storage_bytes(p.k_storage);
}

ffdas_error_t launch_relayout_fi(ffdas_context &h, const DasProblem &p, const void *x, void *dst) {
    hipStream_t st = h.stream;
    hipError_t e = hipErrorInvalidValue;
    using bf = uint16_t;
#define RL(SS, DST) e = p.cplx ? launch_relayout_t<SS, DST, true>(st, p, x, dst) : launch_relayout_t<SS, DST, false>(st, p, x, dst)
    switch (p.x_storage) {
        case Storage::F32:
            if (p.k_storage == Storage::F16) RL(Storage::F32, _Float16);
            else if (p.k_storage == Storage::BF16) RL(Storage::F32, bf);
            else if (p.k_storage == Storage::F32) RL(Storage::F32, float);
            break;
        case Storage::F16:
            if (p.k_storage == Storage::F16) RL(Storage::F16, _Float16);
            break;
        case Storage::BF16:
            if (p.k_storage == Storage::BF16) RL(Storage::BF16, bf);
            break;
        default: break;
    }
#undef RL
    return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

}  // namespace ffdas
