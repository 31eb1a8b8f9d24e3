// Rank truncation (truncated-SVD reconstruction) kernels for gfx950.
//
// Definition (reference: src/truncate_rank/truncate_rank_impl.cuh:382-451, docs/api/
// truncate_rank.md): x is an (m, n) matrix; with p = min(m, n) and the p x p Gram matrix
// of the shorter side, the output keeps the singular components [start, stop) (descending
// order) and removes the rest.  Here the matrix is seen as two families of vectors:
//   * Gram vectors:       p vectors of length L = max(m, n)  (rows when m <= n, else columns);
//   * projection vectors: L vectors of length p              (columns when m <= n, else rows).
// H[a][b] = sum_k v_a[k] conj(v_b[k]) / n is formed over the Gram vectors; with V the
// eigenvectors of H (ascending eigenvalues, so component i is column p-1-i), every projection
// vector w is replaced by V_k V_k^H w (kept set), or by w - V_d V_d^H w (dropped set) when the
// dropped set is the smaller one -- the same projector, with min(kept, dropped) columns.
// For m > n, H is conj(X^H X), whose eigenvectors are the conjugates of the right singular
// vectors; V_k V_k^H applied to the transposed rows is then exactly x W_k W_k^H.
//
// Kernels (all matrix products on the 16x16x4 MFMA, FP32 or FP64, complex as four real
// products):
//   rank_gram_lds_kernel  FP32 split-K Gram over LDS-staged k-tiles: one workgroup per
//                         k-range computes every tile of the upper triangle;
//   rank_gram_kernel      split-K Gram for the other cases: one wave per (upper-triangle
//                         block, k-range); both write block partials in full, no atomics;
//   rank_gram_reduce      sums the partials in a fixed order, scales by 1/n, writes each
//                         upper entry and its conjugate mirror, column-major out;
//   rank_jacobi_kernel    eigendecomposition of H for p <= 64 in one workgroup (parallel
//                         cyclic Jacobi in LDS); larger p go to rocSOLVER (rank_api.cpp);
//   rank_select_kernel    gathers the chosen eigenvector columns into Qh (r x p, conjugated)
//                         and Qn (p x r), zero padded to 16-multiples;
//   rank_project_kernel   per wave 16 projection vectors: C = Qh w (MFMA, k over p), then
//                         out = base +/- Qn C (MFMA, k over r), C kept in registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>

#include "internal.h"

namespace ffdas {
namespace {

template <typename T>
struct Mfma;

// 16x16x4 MFMA: lane (jl, g) = (lane % 16, lane / 16) supplies A[jl][g] and B[g][jl] and
// holds output register v of row row(g, v), column jl: 4g + v for FP32, 4v + g for FP64
// (measured on gfx950).
template <>
struct Mfma<float> {
    typedef float V4 __attribute__((ext_vector_type(4)));
    static constexpr bool kRowsInRegs = true;   // a lane's four outputs are consecutive rows
    static __device__ __forceinline__ int row(int g, int v) { return 4 * g + v; }
    static __device__ __forceinline__ V4 mma(float a, float b, V4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma<double> {
    typedef double V4 __attribute__((ext_vector_type(4)));
    static constexpr bool kRowsInRegs = false;  // rows g, g + 4, g + 8, g + 12
    static __device__ __forceinline__ int row(int g, int v) { return 4 * v + g; }
    static __device__ __forceinline__ V4 mma(double a, double b, V4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// Four elements e, e+se, e+2se, e+3se of one vector starting at element offset `off`
// (complex: interleaved pairs).  `n` of them are valid (the rest read as 0); `vec` selects
// 16-byte loads when se == 1 and the run is 16-byte aligned (checked by the host).
template <typename T, bool CPLX>
__device__ __forceinline__ void load4(const T *__restrict__ x, int64_t off, int64_t se, int n,
                                      bool vec, T (&re)[4], T (&im)[4]) {
    // values go through scalars and are assigned once at the end: stores into the arrays from
    // both branches would be merged into one store with a selected address, which forces the
    // arrays into scratch memory
    T r0 = 0, r1 = 0, r2 = 0, r3 = 0, i0 = 0, i1 = 0, i2 = 0, i3 = 0;
    if (vec && n == 4) {
        if constexpr (sizeof(T) == 4 && !CPLX) {
            const float4 v = *reinterpret_cast<const float4 *>(x + off);
            r0 = v.x; r1 = v.y; r2 = v.z; r3 = v.w;
        } else if constexpr (sizeof(T) == 4 && CPLX) {
            const float4 *q = reinterpret_cast<const float4 *>(x + 2 * off);
            const float4 a = q[0], b = q[1];
            r0 = a.x; i0 = a.y; r1 = a.z; i1 = a.w;
            r2 = b.x; i2 = b.y; r3 = b.z; i3 = b.w;
        } else if constexpr (!CPLX) {
            const double2 *q = reinterpret_cast<const double2 *>(x + off);
            const double2 a = q[0], b = q[1];
            r0 = a.x; r1 = a.y; r2 = b.x; r3 = b.y;
        } else {
            const double2 *q = reinterpret_cast<const double2 *>(x + 2 * off);
            const double2 a = q[0], b = q[1], c = q[2], d = q[3];
            r0 = a.x; i0 = a.y; r1 = b.x; i1 = b.y;
            r2 = c.x; i2 = c.y; r3 = d.x; i3 = d.y;
        }
    } else {
        const int64_t c = CPLX ? 2 : 1;
        if (n > 0) {
            r0 = x[c * off];
            if constexpr (CPLX) i0 = x[c * off + 1];
        }
        if (n > 1) {
            r1 = x[c * (off + se)];
            if constexpr (CPLX) i1 = x[c * (off + se) + 1];
        }
        if (n > 2) {
            r2 = x[c * (off + 2 * se)];
            if constexpr (CPLX) i2 = x[c * (off + 2 * se) + 1];
        }
        if (n > 3) {
            r3 = x[c * (off + 3 * se)];
            if constexpr (CPLX) i3 = x[c * (off + 3 * se) + 1];
        }
    }
    re[0] = r0; re[1] = r1; re[2] = r2; re[3] = r3;
    im[0] = i0; im[1] = i1; im[2] = i2; im[3] = i3;
}

// Four consecutive entries of a zero-padded row-major coefficient matrix (always in bounds).
template <typename T, bool CPLX>
__device__ __forceinline__ void coef4(const T *__restrict__ q, int64_t off, T (&re)[4],
                                      T (&im)[4]) {
    load4<T, CPLX>(q, off, 1, 4, true, re, im);
}

// ---------------------------------------------------------------------------------------
// Gram matrix
// ---------------------------------------------------------------------------------------
// One wave: block (bi, bj), bi <= bj, of size BS = 16*WT, over k in [k_lo, k_hi).  Lane
// (jl, g) = (lane % 16, lane / 16) feeds, at MFMA step s, row jl of a 16-row tile and k-slot
// g, i.e. element k = kb + 4g + s; both operands use the same k mapping, so the sum over k is
// exact whatever the order.  Output D[row(g, v)][jl] of tile (ta, tb) is H-partial
// [16ta + row(g, v)][16tb + jl] of the block.
template <typename T, bool CPLX, int WT>
__global__ __launch_bounds__(64) void rank_gram_kernel(const T *__restrict__ x, int64_t sa,
                                                       int64_t sk, int p, int64_t L,
                                                       int64_t kchunk, int nb, bool vec,
                                                       T *__restrict__ part) {
    using M = Mfma<T>;
    using V4 = typename M::V4;
    constexpr int BS = 16 * WT;
    // Workgroups are dealt to the eight XCDs in turn (id % 8); all blocks of one k-range are
    // given ids on the same XCD, so the rows they share are fetched into one L2 once.
    // splits is a multiple of 8.
    const int nblk = nb * (nb + 1) / 2;
    const int wid = blockIdx.x, xcd = wid & 7, wi = wid >> 3;
    const int split = xcd + 8 * (wi / nblk);
    int idx = wi % nblk, bi = 0;
    while (idx >= nb - bi) {
        idx -= nb - bi;
        bi++;
    }
    const int bj = bi + idx;
    const bool diag = bi == bj;
    const int lane = threadIdx.x, jl = lane & 15, g = lane >> 4;
    const int64_t k_lo = (int64_t)split * kchunk;
    const int64_t k_hi = k_lo + kchunk < L ? k_lo + kchunk : L;   // empty when k_lo >= L

    V4 accr[WT][WT], acci[WT][WT];
#pragma unroll
    for (int a = 0; a < WT; a++)
#pragma unroll
        for (int b = 0; b < WT; b++) {
            accr[a][b] = V4{0, 0, 0, 0};
            acci[a][b] = V4{0, 0, 0, 0};
        }
    int64_t offa[WT], offb[WT];
    bool oka[WT], okb[WT];
#pragma unroll
    for (int t = 0; t < WT; t++) {
        const int ra = bi * BS + 16 * t + jl, rb = bj * BS + 16 * t + jl;
        oka[t] = ra < p;
        okb[t] = rb < p;
        offa[t] = (int64_t)ra * sa;
        offb[t] = (int64_t)rb * sa;
    }

    for (int64_t kb = k_lo; kb < k_hi; kb += 16) {
        const int64_t k0 = kb + 4 * g;
        const int64_t left = k_hi - k0;
        const int nk = left >= 4 ? 4 : (left > 0 ? (int)left : 0);
        T ar[WT][4], ai[WT][4], br[WT][4], bim[WT][4];
#pragma unroll
        for (int t = 0; t < WT; t++)
            load4<T, CPLX>(x, offa[t] + k0 * sk, sk, oka[t] ? nk : 0, vec, ar[t], ai[t]);
        if (diag) {
#pragma unroll
            for (int t = 0; t < WT; t++)
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    br[t][s] = ar[t][s];
                    bim[t][s] = ai[t][s];
                }
        } else {
#pragma unroll
            for (int t = 0; t < WT; t++)
                load4<T, CPLX>(x, offb[t] + k0 * sk, sk, okb[t] ? nk : 0, vec, br[t], bim[t]);
        }
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int a = 0; a < WT; a++)
#pragma unroll
                for (int b = 0; b < WT; b++) {
                    // v_a conj(v_b) = (ar br + ai bi) + i (ai br - ar bi)
                    accr[a][b] = M::mma(ar[a][s], br[b][s], accr[a][b]);
                    if constexpr (CPLX) {
                        accr[a][b] = M::mma(ai[a][s], bim[b][s], accr[a][b]);
                        acci[a][b] = M::mma(ai[a][s], br[b][s], acci[a][b]);
                        acci[a][b] = M::mma(-ar[a][s], bim[b][s], acci[a][b]);
                    }
                }
    }

    constexpr int C = CPLX ? 2 : 1;
    T *dst = part + ((int64_t)split * nblk + (wi % nblk)) * (int64_t)(BS * BS * C);
#pragma unroll
    for (int a = 0; a < WT; a++)
#pragma unroll
        for (int b = 0; b < WT; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int e = (16 * a + M::row(g, v)) * BS + 16 * b + jl;
                if constexpr (CPLX) {
                    dst[2 * e] = accr[a][b][v];
                    dst[2 * e + 1] = acci[a][b][v];
                } else {
                    dst[e] = accr[a][b][v];
                }
            }
}

// Gram over LDS-staged k-tiles (FP32, contiguous 16-byte aligned Gram vectors, p <= 256
// real / 208 complex):
// one workgroup of 8 waves per k-range stages X[0:pp, k:k+32] into LDS (double-buffered,
// each element read from HBM once) and its waves compute every 16 x 16 tile (ta <= tb) of the
// upper triangle, wave w taking tiles w, w + 8, ...  Operands come from LDS as 16-byte reads
// of four consecutive k (the lane's k-slot g covers k = 4g..4g+3 of each 16-k group), so
// both operands use the same k mapping.  Partials: part[split][tile][16][16], the layout of
// the 16 x 16 block kernel, summed by rank_gram_reduce_kernel.
constexpr int kGramKT = 32;      // k per stage
constexpr int kGramWaves = 8;

// MAXT: tiles per wave, ceil(nt (nt + 1) / 2 / 8) for nt = pp / 16 (17 at p = 256)
template <bool CPLX, int MAXT>
__global__ __launch_bounds__(512) void rank_gram_lds_kernel(const float *__restrict__ x,
                                                            int64_t sa, int p, int64_t L,
                                                            int64_t kchunk,
                                                            float *__restrict__ part) {
    using M = Mfma<float>;
    using V4 = M::V4;
    constexpr int C = CPLX ? 2 : 1;
    constexpr int ROWF = kGramKT * C + 4;          // floats per staged row (+4: bank offset)
    extern __shared__ __align__(16) float glds[];
    const int pp = (p + 15) & ~15, nt = pp >> 4, ntri = nt * (nt + 1) / 2;
    float *stage0 = glds, *stage1 = glds + pp * ROWF;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int jl = lane & 15, g = lane >> 4;
    const int64_t k_lo = (int64_t)blockIdx.x * kchunk;
    const int64_t k_hi = k_lo + kchunk < L ? k_lo + kchunk : L;
    const int iters = k_hi > k_lo ? (int)((k_hi - k_lo + kGramKT - 1) / kGramKT) : 0;

    // this wave's tiles, (ta << 8) | tb
    int tab[MAXT];
    int ntiles = 0;
#pragma unroll
    for (int i = 0; i < MAXT; i++) {
        const int t = wave + kGramWaves * i;
        int a = 0, idx = t;
        while (a < nt && idx >= nt - a) {
            idx -= nt - a;
            a++;
        }
        tab[i] = (a << 8) | (a + idx);
        if (t < ntri) ntiles = i + 1;
    }
    V4 accr[MAXT], acci[MAXT];
#pragma unroll
    for (int i = 0; i < MAXT; i++) {
        accr[i] = V4{0, 0, 0, 0};
        acci[i] = V4{0, 0, 0, 0};
    }

    // staging: per row kGramKT * C floats = 8 * C float4; thread -> (row, float4 slot)
    constexpr int F4 = kGramKT * C / 4;
    const int nslots = pp * F4;
    auto fetch = [&](int64_t k0, int slot, float4 &v) {
        const int row = slot / F4, q = slot - row * F4;
        const int64_t k = k0 + (int64_t)q * 4 / C;            // first k of this float4
        v = float4{0, 0, 0, 0};
        if (row >= p) return;
        const float *src = x + ((int64_t)row * sa + k) * C;
        if (k + 4 / C <= k_hi) {
            v = *reinterpret_cast<const float4 *>(src);
        } else {
            if (k < k_hi) v.x = src[0];
            if (k + 1 / C < k_hi) v.y = src[1];
            if (k + 2 / C < k_hi) v.z = src[2];
            if (k + 3 / C < k_hi) v.w = src[3];
        }
    };
    auto put = [&](float *stg, int slot, const float4 &v) {
        const int row = slot / F4, q = slot - row * F4;
        *reinterpret_cast<float4 *>(stg + row * ROWF + q * 4) = v;
    };
    constexpr int kMaxSlotsPerThread = (256 * (kGramKT * 2 / 4) + 511) / 512;   // 8
    float4 pre[kMaxSlotsPerThread];
    if (iters > 0) {
#pragma unroll
        for (int i = 0; i < kMaxSlotsPerThread; i++) {
            const int slot = tid + 512 * i;
            if (slot < nslots) {
                float4 v;
                fetch(k_lo, slot, v);
                put(stage0, slot, v);
            }
        }
    }
    __syncthreads();
    for (int it = 0; it < iters; it++) {
        float *cur = (it & 1) ? stage1 : stage0;
        float *nxt = (it & 1) ? stage0 : stage1;
        const bool more = it + 1 < iters;
        if (more) {
#pragma unroll
            for (int i = 0; i < kMaxSlotsPerThread; i++) {
                const int slot = tid + 512 * i;
                if (slot < nslots) fetch(k_lo + (int64_t)(it + 1) * kGramKT, slot, pre[i]);
            }
        }
#pragma unroll
        for (int kk = 0; kk < kGramKT; kk += 16) {
            // lane's k = kk + 4g + s, s = 0..3
#pragma unroll
            for (int i = 0; i < MAXT; i++) {
                if (i >= ntiles) continue;
                const float *ra = cur + ((tab[i] >> 8) * 16 + jl) * ROWF + (kk + 4 * g) * C;
                const float *rb = cur + ((tab[i] & 255) * 16 + jl) * ROWF + (kk + 4 * g) * C;
                if constexpr (!CPLX) {
                    const float4 a = *reinterpret_cast<const float4 *>(ra);
                    const float4 b = *reinterpret_cast<const float4 *>(rb);
                    accr[i] = M::mma(a.x, b.x, accr[i]);
                    accr[i] = M::mma(a.y, b.y, accr[i]);
                    accr[i] = M::mma(a.z, b.z, accr[i]);
                    accr[i] = M::mma(a.w, b.w, accr[i]);
                } else {
                    const float4 a0 = *reinterpret_cast<const float4 *>(ra);
                    const float4 a1 = *reinterpret_cast<const float4 *>(ra + 4);
                    const float4 b0 = *reinterpret_cast<const float4 *>(rb);
                    const float4 b1 = *reinterpret_cast<const float4 *>(rb + 4);
                    const float ar[4] = {a0.x, a0.z, a1.x, a1.z}, ai[4] = {a0.y, a0.w, a1.y, a1.w};
                    const float br[4] = {b0.x, b0.z, b1.x, b1.z}, bi[4] = {b0.y, b0.w, b1.y, b1.w};
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        // v_a conj(v_b) = (ar br + ai bi) + i (ai br - ar bi)
                        accr[i] = M::mma(ar[s], br[s], accr[i]);
                        accr[i] = M::mma(ai[s], bi[s], accr[i]);
                        acci[i] = M::mma(ai[s], br[s], acci[i]);
                        acci[i] = M::mma(-ar[s], bi[s], acci[i]);
                    }
                }
            }
        }
        if (more) {
#pragma unroll
            for (int i = 0; i < kMaxSlotsPerThread; i++) {
                const int slot = tid + 512 * i;
                if (slot < nslots) put(nxt, slot, pre[i]);
            }
        }
        __syncthreads();
    }
    // partials of this k-range: tile t = wave + 8 i, element (4g + v, jl)
#pragma unroll
    for (int i = 0; i < MAXT; i++) {
        if (i >= ntiles) continue;
        const int t = wave + kGramWaves * i;
        float *dst = part + ((int64_t)blockIdx.x * ntri + t) * 256 * C;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int e = (4 * g + v) * 16 + jl;
            if constexpr (CPLX) {
                dst[2 * e] = accr[i][v];
                dst[2 * e + 1] = acci[i][v];
            } else {
                dst[e] = accr[i][v];
            }
        }
    }
}

size_t gram_lds_bytes(int p, bool cplx) {
    const int pp = (p + 15) & ~15;
    return 2 * (size_t)pp * (kGramKT * (cplx ? 2 : 1) + 4) * sizeof(float);
}

// H (column-major p x p) from the block partials.  Lanes run over the entries of the stored
// blocks in storage order (coalesced reads); 1024 threads = 16 waves per 64 entries, wave w
// sums the splits [w*S/16, (w+1)*S/16) in order, then the 16 sums are added in order.  Each
// upper entry (a <= b) is written twice, as H[a][b] and as its conjugate H[b][a].
template <typename T, bool CPLX>
__global__ __launch_bounds__(1024) void rank_gram_reduce_kernel(const T *__restrict__ part,
                                                                int p, int BS, int nb, int nblk,
                                                                int splits, T scale,
                                                                T *__restrict__ H) {
    __shared__ T red[16][64][2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t blkel = (int64_t)BS * BS;
    const int64_t e = (int64_t)blockIdx.x * 64 + lane;
    bool ok = e < nblk * blkel;
    int a = 0, b = 0;
    if (ok) {
        int idx = (int)(e / blkel), bi = 0;
        while (idx >= nb - bi) {
            idx -= nb - bi;
            bi++;
        }
        const int bj = bi + idx;
        const int le = (int)(e % blkel);
        a = bi * BS + le / BS;
        b = bj * BS + le % BS;
        ok = a <= b && b < p;
    }
    T sr = 0, si = 0;
    if (ok) {
        const int s0 = (int)((int64_t)splits * w / 16), s1 = (int)((int64_t)splits * (w + 1) / 16);
        for (int s = s0; s < s1; s++) {
            const int64_t q = (int64_t)s * nblk * blkel + e;
            if constexpr (CPLX) {
                sr += part[2 * q];
                si += part[2 * q + 1];
            } else {
                sr += part[q];
            }
        }
    }
    red[w][lane][0] = sr;
    red[w][lane][1] = si;
    __syncthreads();
    if (w != 0 || !ok) return;
    T tr = 0, ti = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        tr += red[k][lane][0];
        ti += red[k][lane][1];
    }
    tr *= scale;
    ti = a == b ? T(0) : ti * scale;
    const int64_t up = (int64_t)b * p + a, lo = (int64_t)a * p + b;   // (a, b) and (b, a)
    if constexpr (CPLX) {
        H[2 * up] = tr;
        H[2 * up + 1] = ti;
        H[2 * lo] = tr;
        H[2 * lo + 1] = -ti;
    } else {
        H[up] = tr;
        H[lo] = tr;
    }
}

// Qh[q][f] = conj(V[f][col(q)]) (rp x pp), Qn[f][q] = V[f][col(q)] (pp x rp); V column-major
// p x p.  col(q) = c_first + q for q < q_split, c_second + (q - q_split) otherwise.
template <typename T, bool CPLX>
__global__ void rank_select_kernel(const T *__restrict__ V, int p, int pp, int r, int rp,
                                   int q_split, int c_first, int c_second, T *__restrict__ Qh,
                                   T *__restrict__ Qn) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rp * pp) return;
    const int q = (int)(i / pp), f = (int)(i % pp);
    T vr = 0, vi = 0;
    if (q < r && f < p) {
        const int col = q < q_split ? c_first + q : c_second + (q - q_split);
        const int64_t e = (int64_t)col * p + f;
        vr = CPLX ? V[2 * e] : V[e];
        vi = CPLX ? V[2 * e + 1] : T(0);
    }
    const int64_t h = (int64_t)q * pp + f, n = (int64_t)f * rp + q;
    if constexpr (CPLX) {
        Qh[2 * h] = vr;
        Qh[2 * h + 1] = -vi;
        Qn[2 * n] = vr;
        Qn[2 * n + 1] = vi;
    } else {
        Qh[h] = vr;
        Qn[n] = vr;
    }
}

// ---------------------------------------------------------------------------------------
// Eigendecomposition of a small Hermitian matrix: parallel cyclic Jacobi in one workgroup
// ---------------------------------------------------------------------------------------
// A (pe x pe, pe = p rounded up to even, upper triangle, complex as pairs) lives in LDS, V in
// registers.  Each round pairs the indices by the round-robin tournament (pe - 1 rounds visit
// every pair once); the pe/2 rotations of a round are disjoint, so every 2x2 block (P, Q) of
// A is updated independently as G_P^H B G_Q, and V's rows as V G, with G the unitary Schur
// rotation of the pair (its form is given at the kernel).  A rotation is skipped when
// |a_ij| <= eps * ||A||_F, and the iteration ends after the sweep that leaves no larger
// off-diagonal element.  Output: eigenvectors in the columns of H (column-major), ascending
// eigenvalues, as the LAPACK-style solver returns them; info = 1 if max_sweeps did not
// converge.
template <typename T>
struct Eps;
template <>
struct Eps<float> {
    static constexpr float v = 1.1920929e-07f;
};
template <>
struct Eps<double> {
    static constexpr double v = 2.220446049250313e-16;
};

// A and V are kept row-major with a row stride of pe + 1 elements, so a column of either
// (the Hermitian mirror stores, the diagonal) falls on distinct LDS banks
size_t jacobi_lds_bytes(int pe, size_t esz, size_t tsz) {
    const int npair = pe / 2;
    const size_t ntri = (size_t)npair * (npair + 1) / 2;
    return 2 * (size_t)pe * (pe + 1) * esz + (size_t)npair * 4 * tsz +
           ntri * sizeof(unsigned short) + 64;
}

// rotation arithmetic: hardware reciprocal, square root and reciprocal square root (about
// 1 ulp) in FP32, IEEE operations in FP64
__device__ __forceinline__ float jdiv(float a, float b) { return __fdividef(a, b); }
__device__ __forceinline__ double jdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float jrsqrt(float a) { return rsqrtf(a); }
__device__ __forceinline__ double jrsqrt(double a) { return 1.0 / sqrt(a); }
__device__ __forceinline__ float jsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double jsqrt(double a) { return sqrt(a); }

// the pair P of round `round` of the round-robin tournament over pe (even) indices: slot 0
// holds index 0, slots 1 .. pe-1 hold 1 + (slot - 1 + round) mod (pe - 1); pair P takes
// slots P and pe-1-P; returned as i < j
__device__ __forceinline__ int2 jslots(int P, int round, int pe) {
    int a = P - 1 + round, b = pe - 2 - P + round;   // both below 2 (pe - 1)
    if (a >= pe - 1) a -= pe - 1;
    if (b >= pe - 1) b -= pe - 1;
    return int2{P == 0 ? 0 : 1 + a, 1 + b};   // indices in slots P and pe-1-P
}
__device__ __forceinline__ int2 jpair(int P, int round, int pe) {
    const int2 s = jslots(P, round, pe);
    return s.x < s.y ? s : int2{s.y, s.x};
}

template <typename T>
struct JVec;
template <>
struct JVec<float> {
    using V2 = float2;
    using V4 = float4;
};
template <>
struct JVec<double> {
    using V2 = double2;
    using V4 = double4;
};

// LDS element (r, i) of a complex or real matrix: one 8/16-byte access per element
template <typename T, bool CPLX>
__device__ __forceinline__ void jld(const T *M, int e, T &re, T &im) {
    if constexpr (CPLX) {
        const typename JVec<T>::V2 v = reinterpret_cast<const typename JVec<T>::V2 *>(M)[e];
        re = v.x;
        im = v.y;
    } else {
        re = M[e];
        im = T(0);
    }
}
template <typename T, bool CPLX>
__device__ __forceinline__ void jst(T *M, int e, T re, T im) {
    if constexpr (CPLX) {
        typename JVec<T>::V2 v;
        v.x = re;
        v.y = im;
        reinterpret_cast<typename JVec<T>::V2 *>(M)[e] = v;
    } else {
        M[e] = re;
    }
}

// Upper-triangle storage: element (r, c) of the Hermitian A lives at U[min][max], conjugated
// when r > c; the diagonal block of a pair is written as a real diagonal and an exact zero.
template <typename T, bool CPLX>
__device__ __forceinline__ void jget(const T *A, int ld, int r, int c, T &re, T &im) {
    const bool up = r <= c;   // one access either way: address and sign are selected
    jld<T, CPLX>(A, up ? r * ld + c : c * ld + r, re, im);
    if constexpr (CPLX) im = up ? im : -im;
}
template <typename T, bool CPLX>
__device__ __forceinline__ void jput(T *A, int ld, int r, int c, T re, T im) {
    const bool up = r < c;
    jst<T, CPLX>(A, up ? r * ld + c : c * ld + r, re, up ? im : -im);
}

// One-lane moves across the whole wave on the VALU (DPP): jnext gives lane l the value of
// lane l + 1 (wave_shl:1), jprev that of lane l - 1 (wave_shr:1); the lane without a source
// keeps its own value.
__device__ __forceinline__ int jdpp_i(int v, bool next) {
    return next ? __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false)
                : __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ float jdpp(float v, bool next) {
    return __int_as_float(jdpp_i(__float_as_int(v), next));
}
__device__ __forceinline__ double jdpp(double v, bool next) {
    const long long b = __double_as_longlong(v);
    const int lo = jdpp_i((int)(unsigned)b, next), hi = jdpp_i((int)(b >> 32), next);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// V rows per thread at 1024 threads: 2 with 32-lane row slots (pe <= 64); with 64-lane slots
// (16 row slots) the most pe whose A and V fit in 160 KB of LDS allows: 138 (4-byte
// elements, 9 rows), 98 (8-byte, 7 rows), 70 (16-byte, 5 rows)
constexpr int jacobi_vrows(int sw, size_t esz) {
    return sw == 32 ? 2 : (esz >= 16 ? 5 : (esz >= 8 ? 7 : 9));
}

// SW: lanes per V row slot (32 for pe <= 64, 64 above); 1024 threads.
// Each round uses G = [[c, sigma], [-conj(sigma), c]] with sigma = s e^{i phi}, a_ij = r e^{i phi}:
// G^H [[a_ii, a_ij], [conj(a_ij), a_jj]] G is diagonal for t = s / c = 2 r sgn(d) / (|d| + h),
// d = a_jj - a_ii, h = sqrt(d^2 + 4 r^2), so c = 1 / sqrt(1 + 4 r^2 / (|d| + h)^2) and
// sigma = 2 sgn(d) c a_ij / (|d| + h): no square root of |a_ij| and no phase division.
// (Real data: the same with sigma = s, a_ij signed.)
template <typename T, bool CPLX, int SW>
__global__ __launch_bounds__(1024) void rank_jacobi_kernel(T *H, int p, int pe, int max_sweeps,
                                                           int *info) {
    using V4 = typename JVec<T>::V4;
    constexpr int VR = jacobi_vrows(SW, sizeof(T) * (CPLX ? 2 : 1));   // >= ceil(pe / (1024 / SW))
    extern __shared__ __align__(16) unsigned char jlds[];
    constexpr int C = CPLX ? 2 : 1;
    const int npair = pe / 2, ld = pe + 1;
    T *A = reinterpret_cast<T *>(jlds);
    T *V = A + (size_t)pe * ld * C;
    // per pair: (c, Re sigma, Im sigma, 0), one vector access
    V4 *prm = reinterpret_cast<V4 *>(V + (size_t)pe * ld * C);
    // the blocks (P, Q), P <= Q, of a round: P << 8 | Q
    unsigned short *tri = reinterpret_cast<unsigned short *>(prm + npair);
    const int ntri = npair * (npair + 1) / 2;
    __shared__ double red[16];
    const int tid = threadIdx.x, nt = blockDim.x;
    // V stays in registers: lane k = tid % SW of row slot rs = tid / SW holds, for rows
    // rs + v * nrs, the elements of its pair's two tournament slots, a = slot k and
    // b = slot pe-1-k; after each round the values move with the slots' indices (slot m
    // takes slot m+1's index, slot 0 keeps index 0), which is a one-lane shift.
    const int k = tid % SW, rs = tid / SW, nrs = nt / SW;
    const bool vown = k < npair;

    // ||H||_F in FP64; A = H 2^-e with ||A||_F near 1.  Power-of-two scaling is exact and the
    // rotations do not depend on scale, so d^2 + 4 r^2 stays far from overflow
    double sq = 0;
    for (int e = tid; e < p * p; e += nt) {
        const double ar = CPLX ? H[2 * e] : H[e], ai = CPLX ? H[2 * e + 1] : T(0);
        sq += ar * ar + ai * ai;
    }
    for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o);
    if ((tid & 63) == 0) red[tid >> 6] = sq;
    __syncthreads();
    double norm2 = 0;
    for (int w = 0; w < nt / 64; w++) norm2 += red[w];
    int ex = 0;
    if (norm2 > 0) frexp(sqrt(norm2), &ex);
    ex = max(-126, min(126, ex));
    const T sc = (T)ldexp(1.0, -ex);
    const T thresh = (T)(Eps<T>::v * sqrt(norm2) * ldexp(1.0, -ex));
    const T thresh2 = thresh * thresh;

    // load A (H is column-major: (r, c) at c * p + r)
    for (int e = tid; e < pe * pe; e += nt) {
        const int r = e / pe, c = e % pe;
        T ar = 0, ai = 0;
        if (r < p && c < p) {
            const int64_t h = (int64_t)c * p + r;
            ar = (CPLX ? H[2 * h] : H[h]) * sc;
            ai = CPLX ? H[2 * h + 1] * sc : T(0);
        }
        jst<T, CPLX>(A, r * ld + c, ar, ai);
    }
    int myPQ = 0;   // this thread's first block, kept in a register (the only one if nt >= ntri)
    for (int w = tid; w < ntri; w += nt) {
        int P = 0, idx = w;
        while (idx >= npair - P) {
            idx -= npair - P;
            P++;
        }
        const int PQ = (P << 8) | (P + idx);
        tri[w] = (unsigned short)PQ;
        if (w == tid) myPQ = PQ;
    }
    // V = I: slot k holds index k, slot pe-1-k index pe-1-k at round 0
    T var[VR], vai[VR], vbr[VR], vbi[VR];
#pragma unroll
    for (int v = 0; v < VR; v++) {
        const int row = rs + v * nrs;
        var[v] = row == k ? T(1) : T(0);
        vbr[v] = row == pe - 1 - k ? T(1) : T(0);
        vai[v] = vbi[v] = T(0);
    }
    __syncthreads();

    // A sweep rotates the pairs whose |a_ij|^2 exceeds thresh^2.  The iteration ends once no
    // off-diagonal element does (checked after each sweep: the next sweep would rotate
    // nothing, so skipping it changes no result).
    auto converged = [&]() {
        int any = 0;
        for (int e = tid; e < p * p; e += nt) {
            const int r = e / p, c = e % p;
            if (r < c) {
                T zr, zi;
                jld<T, CPLX>(A, r * ld + c, zr, zi);
                any |= (CPLX ? zr * zr + zi * zi : zr * zr) > thresh2;
            }
        }
        return __syncthreads_or(any) == 0;
    };
    int sweep = 0;
    bool done = converged();
    for (; sweep < max_sweeps && !done; sweep++) {
        for (int round = 0; round < pe - 1; round++) {
            // rotation parameters of this round's pairs
            if (tid < npair) {
                const int2 ij = jpair(tid, round, pe);
                const int i = ij.x, j = ij.y;
                T c = 1, sr = 0, si = 0;
                if (j < p) {
                    T aii, ajj, zr, zi, dummy;
                    jld<T, CPLX>(A, i * ld + i, aii, dummy);
                    jld<T, CPLX>(A, j * ld + j, ajj, dummy);
                    jld<T, CPLX>(A, i * ld + j, zr, zi);
                    const T r2 = CPLX ? zr * zr + zi * zi : zr * zr;
                    if (r2 > thresh2) {
                        const T d = ajj - aii;
                        const T inv = jdiv(T(1), fabs(d) + jsqrt(d * d + 4 * r2));
                        c = jrsqrt(1 + 4 * r2 * (inv * inv));
                        const T f = (d >= 0 ? T(2) : T(-2)) * c * inv;
                        sr = f * zr;
                        si = CPLX ? f * zi : T(0);
                    }
                }
                V4 q;
                q.x = c;
                q.y = sr;
                q.z = si;
                q.w = 0;
                prm[tid] = q;
            }
            __syncthreads();
            // blocks (P, Q), P <= Q: B' = G_P^H B G_Q in upper-triangle storage
            for (int w = tid; w < ntri; w += nt) {
                const int PQ = w == tid ? myPQ : (int)tri[w], P = PQ >> 8, Q = PQ & 255;
                const int2 ijP = jpair(P, round, pe), ijQ = jpair(Q, round, pe);
                const int iP = ijP.x, jP = ijP.y, iQ = ijQ.x, jQ = ijQ.y;
                const V4 qP = prm[P], qQ = prm[Q];
                T b00r, b01r, b10r, b11r, b00i, b01i, b10i, b11i;
                jget<T, CPLX>(A, ld, iP, iQ, b00r, b00i);
                jget<T, CPLX>(A, ld, iP, jQ, b01r, b01i);
                jget<T, CPLX>(A, ld, jP, iQ, b10r, b10i);
                jget<T, CPLX>(A, ld, jP, jQ, b11r, b11i);
                const T cP = qP.x, sPr = qP.y, sPi = qP.z, cQ = qQ.x, sQr = qQ.y, sQi = qQ.z;
                if (sPr == 0 && sPi == 0 && sQr == 0 && sQi == 0) continue;
                // N = B G_Q: n_x0 = cQ b_x0 - conj(sQ) b_x1, n_x1 = sQ b_x0 + cQ b_x1
                T n00r, n00i, n01r, n01i, n10r, n10i, n11r, n11i;
                if constexpr (CPLX) {
                    n00r = cQ * b00r - (sQr * b01r + sQi * b01i);
                    n00i = cQ * b00i - (sQr * b01i - sQi * b01r);
                    n01r = (sQr * b00r - sQi * b00i) + cQ * b01r;
                    n01i = (sQr * b00i + sQi * b00r) + cQ * b01i;
                    n10r = cQ * b10r - (sQr * b11r + sQi * b11i);
                    n10i = cQ * b10i - (sQr * b11i - sQi * b11r);
                    n11r = (sQr * b10r - sQi * b10i) + cQ * b11r;
                    n11i = (sQr * b10i + sQi * b10r) + cQ * b11i;
                } else {
                    n00r = cQ * b00r - sQr * b01r;
                    n01r = sQr * b00r + cQ * b01r;
                    n10r = cQ * b10r - sQr * b11r;
                    n11r = sQr * b10r + cQ * b11r;
                    n00i = n01i = n10i = n11i = 0;
                }
                // B' = G_P^H N: o_0y = cP n_0y - sP n_1y, o_1y = conj(sP) n_0y + cP n_1y
                T o00r, o00i, o01r, o01i, o10r, o10i, o11r, o11i;
                if constexpr (CPLX) {
                    o00r = cP * n00r - (sPr * n10r - sPi * n10i);
                    o00i = cP * n00i - (sPr * n10i + sPi * n10r);
                    o01r = cP * n01r - (sPr * n11r - sPi * n11i);
                    o01i = cP * n01i - (sPr * n11i + sPi * n11r);
                    o10r = (sPr * n00r + sPi * n00i) + cP * n10r;
                    o10i = (sPr * n00i - sPi * n00r) + cP * n10i;
                    o11r = (sPr * n01r + sPi * n01i) + cP * n11r;
                    o11i = (sPr * n01i - sPi * n01r) + cP * n11i;
                } else {
                    o00r = cP * n00r - sPr * n10r;
                    o01r = cP * n01r - sPr * n11r;
                    o10r = sPr * n00r + cP * n10r;
                    o11r = sPr * n01r + cP * n11r;
                    o00i = o01i = o10i = o11i = 0;
                }
                if (P == Q) {   // the rotated pair itself: real diagonal, exact zero off it
                    jst<T, CPLX>(A, iP * ld + iP, o00r, T(0));
                    jst<T, CPLX>(A, jP * ld + jP, o11r, T(0));
                    jst<T, CPLX>(A, iP * ld + jP, T(0), T(0));
                } else {
                    jput<T, CPLX>(A, ld, iP, iQ, o00r, o00i);
                    jput<T, CPLX>(A, ld, iP, jQ, o01r, o01i);
                    jput<T, CPLX>(A, ld, jP, iQ, o10r, o10i);
                    jput<T, CPLX>(A, ld, jP, jQ, o11r, o11i);
                }
            }
            // V' = V G on the registers: v_i' = c v_i - conj(sigma) v_j, v_j' = sigma v_i + c v_j.
            // With (a, b) the lane's two slots: a' = c a + s1 b, b' = c b + s2 a, where
            // (s1, s2) = (-conj(sigma), sigma) if a holds i, else (sigma, -conj(sigma)).
            if (vown) {
                const V4 q = prm[k];
                if (q.y != 0 || q.z != 0) {
                    const int2 sl = jslots(k, round, pe);
                    const T c = q.x, si = q.z;
                    const T s1r = sl.x < sl.y ? -q.y : q.y, s2r = -s1r;
#pragma unroll
                    for (int v = 0; v < VR; v++) {
                        const T ar = var[v], ai = vai[v], br = vbr[v], bi = vbi[v];
                        if constexpr (CPLX) {
                            var[v] = c * ar + (s1r * br - si * bi);
                            vai[v] = c * ai + (s1r * bi + si * br);
                            vbr[v] = c * br + (s2r * ar - si * ai);
                            vbi[v] = c * bi + (s2r * ai + si * ar);
                        } else {
                            var[v] = c * ar + s1r * br;
                            vbr[v] = c * br + s2r * ar;
                        }
                    }
                }
            }
            // move to the next round's slots: a <- slot k+1 (lane k+1's a; the last pair its
            // own b), b <- slot pe-k (lane k-1's b; pair 0 slot 1, lane 1's a or, for pe = 2,
            // its own b).  Every lane takes part in the moves (all lanes are active here).
#pragma unroll
            for (int v = 0; v < VR; v++) {
                const T dar = jdpp(var[v], true), ubr = jdpp(vbr[v], false);
                const T dai = CPLX ? jdpp(vai[v], true) : T(0);
                const T ubi = CPLX ? jdpp(vbi[v], false) : T(0);
                const T nar = k == 0 ? var[v] : (k == npair - 1 ? vbr[v] : dar);
                const T nai = k == 0 ? vai[v] : (k == npair - 1 ? vbi[v] : dai);
                const T nbr = k == 0 ? (npair > 1 ? dar : vbr[v]) : ubr;
                const T nbi = k == 0 ? (npair > 1 ? dai : vbi[v]) : ubi;
                var[v] = nar;
                vai[v] = nai;
                vbr[v] = nbr;
                vbi[v] = nbi;
            }
            __syncthreads();
        }
        done = converged();
    }
    // every sweep ends where it began: slot k holds index k, slot pe-1-k index pe-1-k
    if (vown) {
#pragma unroll
        for (int v = 0; v < VR; v++) {
            const int row = rs + v * nrs;
            if (row < pe) {
                jst<T, CPLX>(V, row * ld + k, var[v], vai[v]);
                jst<T, CPLX>(V, row * ld + pe - 1 - k, vbr[v], vbi[v]);
            }
        }
    }
    __syncthreads();

    // ascending order of the eigenvalues; ties by index
    for (int w = tid; w < p * p; w += nt) {
        const int row = w % p, k = w / p;
        T dk, dummy;
        jld<T, CPLX>(A, k * ld + k, dk, dummy);
        int rank = 0;
        for (int l = 0; l < p; l++) {
            T dl;
            jld<T, CPLX>(A, l * ld + l, dl, dummy);
            rank += (dl < dk || (dl == dk && l < k)) ? 1 : 0;
        }
        T vr, vi;
        jld<T, CPLX>(V, row * ld + k, vr, vi);
        const int64_t h = (int64_t)rank * p + row;
        H[C * h] = vr;
        if constexpr (CPLX) H[C * h + 1] = vi;
    }
    if (tid == 0) {
        info[0] = done ? 0 : 1;
        info[1] = sweep;   // sweeps run (diagnostics)
    }
}

// ---------------------------------------------------------------------------------------
// Projection
// ---------------------------------------------------------------------------------------
// 256 threads = 4 waves, each wave 16 projection vectors j = j0 + jl.  Coefficient tiles
// q0 + [0, 16*RT).  base: 0 -> out = sign*P w, 1 -> out = x + sign*P w, 2 -> out += sign*P w.
// NK > 0 (FP32, pp <= 16*NK): the x values a lane loads in the first product are exactly the
// ones it needs again for base 1 (rows 4g + s of each 16-block), so they stay in registers
// and x is read from memory once.
template <typename T, bool CPLX, int RT, int NK>
__global__ __launch_bounds__(256) void rank_project_kernel(
    const T *x, int64_t xsj, int64_t xsf, bool xvec, T *out, int64_t osj,
    int64_t osf, bool ovec, int64_t L, int p, int pp, int rp, int q0,
    const T *__restrict__ Qh, const T *__restrict__ Qn, int base, T sign) {
    using M = Mfma<T>;
    using V4 = typename M::V4;
    static_assert(NK == 0 || M::kRowsInRegs, "register reuse needs the FP32 output layout");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int jl = lane & 15, g = lane >> 4;
    const int64_t j = ((int64_t)blockIdx.x * 4 + wave) * 16 + jl;
    const bool jok = j < L;
    const int64_t xo = jok ? j * xsj : 0, oo = jok ? j * osj : 0;

    V4 cr[RT], ci[RT];
#pragma unroll
    for (int t = 0; t < RT; t++) {
        cr[t] = V4{0, 0, 0, 0};
        ci[t] = V4{0, 0, 0, 0};
    }
    // C[q][j] = sum_f Qh[q][f] w_j[f]
    auto product1 = [&](int kb, T (&wr)[4], T (&wi)[4]) {
        const int f0 = kb + 4 * g;
        const int nf = !jok ? 0 : (p - f0 >= 4 ? 4 : (p - f0 > 0 ? p - f0 : 0));
        load4<T, CPLX>(x, xo + (int64_t)f0 * xsf, xsf, nf, xvec, wr, wi);
#pragma unroll
        for (int t = 0; t < RT; t++) {
            T hr[4], hi[4];
            coef4<T, CPLX>(Qh, (int64_t)(q0 + 16 * t + jl) * pp + f0, hr, hi);
#pragma unroll
            for (int s = 0; s < 4; s++) {
                cr[t] = M::mma(hr[s], wr[s], cr[t]);
                if constexpr (CPLX) {
                    cr[t] = M::mma(-hi[s], wi[s], cr[t]);
                    ci[t] = M::mma(hr[s], wi[s], ci[t]);
                    ci[t] = M::mma(hi[s], wr[s], ci[t]);
                }
            }
        }
    };
    T kr[NK > 0 ? NK : 1][4], ki[NK > 0 ? NK : 1][4];
    if constexpr (NK > 0) {
#pragma unroll
        for (int b = 0; b < NK; b++)
            if (16 * b < pp) product1(16 * b, kr[b], ki[b]);
    } else {
        for (int kb = 0; kb < pp; kb += 16) {
            T wr[4], wi[4];
            product1(kb, wr, wi);
        }
    }
    // D[f][j] = sum_q Qn[f][q] C[q][j]; lane holds C[16t + row(g, s)][jl] = c[t][s], the B
    // operand of k-slot g at step s when the A operand uses the same q = 16t + row(g, s).
    // xr/xi: the lane's x values of this 16-block when held in registers.
    auto product2 = [&](int ft, const T (&xr)[4], const T (&xi)[4], bool held) {
        V4 dr = V4{0, 0, 0, 0}, di = V4{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < RT; t++) {
            T nr[4], ni[4];
            if constexpr (M::kRowsInRegs)
                coef4<T, CPLX>(Qn, (int64_t)(ft + jl) * rp + q0 + 16 * t + 4 * g, nr, ni);
            else
                load4<T, CPLX>(Qn, (int64_t)(ft + jl) * rp + q0 + 16 * t + g, 4, 4, false, nr,
                               ni);
#pragma unroll
            for (int s = 0; s < 4; s++) {
                dr = M::mma(nr[s], cr[t][s], dr);
                if constexpr (CPLX) {
                    dr = M::mma(-ni[s], ci[t][s], dr);
                    di = M::mma(nr[s], ci[t][s], di);
                    di = M::mma(ni[s], cr[t][s], di);
                }
            }
        }
        if constexpr (!M::kRowsInRegs) {
            // lane holds D[ft + g + 4v][j]
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int f = ft + g + 4 * v;
                if (!jok || f >= p) continue;
                const int64_t xe = xo + (int64_t)f * xsf, oe = oo + (int64_t)f * osf;
                T br = 0, bi = 0;
                if (base == 1) {
                    br = CPLX ? x[2 * xe] : x[xe];
                    bi = CPLX ? x[2 * xe + 1] : T(0);
                } else if (base == 2) {
                    br = CPLX ? out[2 * oe] : out[oe];
                    bi = CPLX ? out[2 * oe + 1] : T(0);
                }
                if constexpr (CPLX) {
                    out[2 * oe] = br + sign * dr[v];
                    out[2 * oe + 1] = bi + sign * di[v];
                } else {
                    out[oe] = br + sign * dr[v];
                }
            }
            return;
        }
        // lane holds D[ft + 4g + v][j]
        const int f0 = ft + 4 * g;
        const int nf = !jok ? 0 : (p - f0 >= 4 ? 4 : (p - f0 > 0 ? p - f0 : 0));
        if (nf == 0) return;
        T br[4], bi[4];
        if (base == 1 && held) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                br[v] = xr[v];
                bi[v] = xi[v];
            }
        } else if (base == 1) {
            load4<T, CPLX>(x, xo + (int64_t)f0 * xsf, xsf, nf, xvec, br, bi);
        } else if (base == 2) {
            load4<T, CPLX>(out, oo + (int64_t)f0 * osf, osf, nf, ovec, br, bi);
        } else {
#pragma unroll
            for (int v = 0; v < 4; v++) br[v] = bi[v] = T(0);
        }
        T yr[4], yi[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            yr[v] = br[v] + sign * dr[v];
            yi[v] = bi[v] + sign * di[v];
        }
        const int64_t e0 = oo + (int64_t)f0 * osf;
        if (ovec && nf == 4) {
            if constexpr (sizeof(T) == 4 && !CPLX) {
                *reinterpret_cast<float4 *>(out + e0) = float4{yr[0], yr[1], yr[2], yr[3]};
            } else if constexpr (sizeof(T) == 4 && CPLX) {
                float4 *q = reinterpret_cast<float4 *>(out + 2 * e0);
                q[0] = float4{yr[0], yi[0], yr[1], yi[1]};
                q[1] = float4{yr[2], yi[2], yr[3], yi[3]};
            } else if constexpr (!CPLX) {
                double2 *q = reinterpret_cast<double2 *>(out + e0);
                q[0] = double2{yr[0], yr[1]};
                q[1] = double2{yr[2], yr[3]};
            } else {
                double2 *q = reinterpret_cast<double2 *>(out + 2 * e0);
#pragma unroll
                for (int v = 0; v < 4; v++) q[v] = double2{yr[v], yi[v]};
            }
        } else {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                if (v < nf) {
                    const int64_t e = e0 + v * osf;
                    if constexpr (CPLX) {
                        out[2 * e] = yr[v];
                        out[2 * e + 1] = yi[v];
                    } else {
                        out[e] = yr[v];
                    }
                }
            }
        }
    };
    if constexpr (NK > 0) {
#pragma unroll
        for (int b = 0; b < NK; b++)
            if (16 * b < pp) product2(16 * b, kr[b], ki[b], true);
    } else {
        const T zr[4] = {0, 0, 0, 0}, zi[4] = {0, 0, 0, 0};
        for (int ft = 0; ft < pp; ft += 16) product2(ft, zr, zi, false);
    }
}

// Strided 2-D copy (keep-all case, or the private copy of x for in-place multi-pass calls).
template <typename T, bool CPLX>
__global__ void rank_copy_kernel(const T *__restrict__ x, int64_t xs0, int64_t xs1, T *out,
                                 int64_t os0, int64_t os1, int64_t m, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m * n) return;
    const int64_t a = i / n, b = i % n;
    const int64_t s = a * xs0 + b * xs1, d = a * os0 + b * os1;
    if constexpr (CPLX) {
        out[2 * d] = x[2 * s];
        out[2 * d + 1] = x[2 * s + 1];
    } else {
        out[d] = x[s];
    }
}

template <typename T, bool CPLX, int WT>
void gram_launch(hipStream_t st, const RankPlan &pl, const void *x, void *part, void *H) {
    const int nblk = pl.nb * (pl.nb + 1) / 2;
    {
        ProfScope ps(st, "rank_gram");
        hipLaunchKernelGGL((rank_gram_kernel<T, CPLX, WT>), dim3(nblk * pl.splits), dim3(64), 0,
                           st, (const T *)x, pl.gsa, pl.gsk, pl.p, pl.L, pl.kchunk, pl.nb,
                           pl.gvec, (T *)part);
    }
    const int64_t ne = (int64_t)nblk * pl.bs * pl.bs;
    ProfScope ps(st, "rank_gram_reduce");
    hipLaunchKernelGGL((rank_gram_reduce_kernel<T, CPLX>), dim3((unsigned)((ne + 63) / 64)),
                       dim3(1024), 0, st, (const T *)part, pl.p, pl.bs, pl.nb, nblk, pl.splits,
                       (T)(1.0 / (double)pl.n), (T *)H);
}

template <typename T, bool CPLX>
void gram_dispatch(hipStream_t st, const RankPlan &pl, const void *x, void *part, void *H) {
    if constexpr (sizeof(T) == 4) {
        if (pl.lds_gram) {
            const int nt = pl.pp / 16, ntri = nt * (nt + 1) / 2;
            const size_t lds = gram_lds_bytes(pl.p, CPLX);
            {
                ProfScope ps(st, "rank_gram");
                const int per_wave = (ntri + kGramWaves - 1) / kGramWaves;
#define FFDAS_GRAM_LDS(MT)                                                                     \
    do {                                                                                      \
        (void)hipFuncSetAttribute((const void *)rank_gram_lds_kernel<CPLX, MT>,               \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);      \
        hipLaunchKernelGGL((rank_gram_lds_kernel<CPLX, MT>), dim3(pl.splits), dim3(512), lds, \
                           st, (const float *)x, pl.gsa, pl.p, pl.L, pl.kchunk,               \
                           (float *)part);                                                    \
    } while (0)
                if (per_wave <= 2)
                    FFDAS_GRAM_LDS(2);
                else if (per_wave <= 5)
                    FFDAS_GRAM_LDS(5);
                else if (per_wave <= 8)
                    FFDAS_GRAM_LDS(8);
                else if (per_wave <= 12)
                    FFDAS_GRAM_LDS(12);
                else
                    FFDAS_GRAM_LDS(17);
#undef FFDAS_GRAM_LDS
            }
            ProfScope ps(st, "rank_gram_reduce");
            const int64_t ne = (int64_t)ntri * 256;
            hipLaunchKernelGGL((rank_gram_reduce_kernel<float, CPLX>),
                               dim3((unsigned)((ne + 63) / 64)), dim3(1024), 0, st,
                               (const float *)part, pl.p, 16, nt, ntri, pl.splits,
                               (float)(1.0 / (double)pl.n), (float *)H);
            return;
        }
    }
    switch (pl.bs) {
        case 16: gram_launch<T, CPLX, 1>(st, pl, x, part, H); break;
        case 32: gram_launch<T, CPLX, 2>(st, pl, x, part, H); break;
        default: gram_launch<T, CPLX, 4>(st, pl, x, part, H); break;
    }
}

template <typename T, bool CPLX, int RT, int NK>
void project_launch_nk(hipStream_t st, const RankPlan &pl, const void *x, void *out, int q0,
                       const void *Qh, const void *Qn, int base, double sign) {
    const unsigned blocks = (unsigned)((pl.L + 63) / 64);
    ProfScope ps(st, "rank_project");
    hipLaunchKernelGGL((rank_project_kernel<T, CPLX, RT, NK>), dim3(blocks), dim3(256), 0, st,
                       (const T *)x, pl.xsj, pl.xsf, pl.xvec, (T *)out, pl.osj, pl.osf, pl.ovec,
                       pl.L, pl.p, pl.pp, pl.rp, q0, (const T *)Qh, (const T *)Qn, base, (T)sign);
}

// FP32 keeps the lane's x values in registers between the two products when they fit: up to
// 4 blocks of 16 (p <= 64) for any r, up to 16 blocks (p <= 256) for r <= 32
template <typename T, bool CPLX, int RT>
void project_launch(hipStream_t st, const RankPlan &pl, const void *x, void *out, int q0,
                    const void *Qh, const void *Qn, int base, double sign) {
    constexpr bool fp32 = sizeof(T) == 4;
    if (fp32 && pl.pp <= 64 && base == 1)
        project_launch_nk<T, CPLX, RT, fp32 ? 4 : 0>(st, pl, x, out, q0, Qh, Qn, base, sign);
    else if (fp32 && RT <= 2 && pl.pp <= 256 && base == 1)
        project_launch_nk<T, CPLX, RT, fp32 && RT <= 2 ? 16 : 0>(st, pl, x, out, q0, Qh, Qn,
                                                                 base, sign);
    else
        project_launch_nk<T, CPLX, RT, 0>(st, pl, x, out, q0, Qh, Qn, base, sign);
}

template <typename T, bool CPLX>
void project_dispatch(hipStream_t st, const RankPlan &pl, const void *x, void *out,
                      const void *Qh, const void *Qn) {
    // passes of at most 8 coefficient tiles; the first writes, the others accumulate
    const double sign = pl.keep ? 1.0 : -1.0;
    for (int pass = 0; pass < pl.passes; pass++) {
        const int q0 = pass * 16 * pl.rt;
        const int base = pass > 0 ? 2 : (pl.keep ? 0 : 1);
        switch (pl.rt) {
            case 1: project_launch<T, CPLX, 1>(st, pl, x, out, q0, Qh, Qn, base, sign); break;
            case 2: project_launch<T, CPLX, 2>(st, pl, x, out, q0, Qh, Qn, base, sign); break;
            case 4: project_launch<T, CPLX, 4>(st, pl, x, out, q0, Qh, Qn, base, sign); break;
            default: project_launch<T, CPLX, 8>(st, pl, x, out, q0, Qh, Qn, base, sign); break;
        }
    }
}

// ---------------------------------------------------------------------------------------
// FP32 eigenvectors (real p <= 224, complex p <= 160): Householder tridiagonalisation in one
// workgroup, the tridiagonal's chosen eigenpairs by bisection and inverse iteration in one
// workgroup, then the back-transformation of those vectors alone.
// ---------------------------------------------------------------------------------------

// DPP lane moves (GFX9 controls): quad permutations, row mirrors, and the row broadcasts of
// lane 15 / lane 31 into the rows that row_mask enables (other rows receive 0).
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dppf(float v) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}

// Sum over each row of 16 lanes, in every lane of the row.
__device__ __forceinline__ float row16_sum(float v) {
    v += dppf<0xb1>(v);    // quad_perm [1, 0, 3, 2]
    v += dppf<0x4e>(v);    // quad_perm [2, 3, 0, 1]
    v += dppf<0x141>(v);   // row_half_mirror
    v += dppf<0x140>(v);   // row_mirror
    return v;
}

// Sum over each half-wave (lanes 0-31 and 32-63), valid in lanes 16-31 and 48-63.
__device__ __forceinline__ float half_wave_sum_hi(float v) {
    v = row16_sum(v);
    return v + dppf<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
}

// Sum over the wave, the same value in every lane.
__device__ __forceinline__ float wave_total(float v) {
    v = half_wave_sum_hi(v);
    v += dppf<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__host__ __device__ constexpr int packed_column(int p, int k) { return k * p - k * (k + 1) / 2; }

// u_j = 1 at j = k + 1, (x_j * s) for k + 1 < j < p, 0 elsewhere (s complex for a Hermitian A).
template <bool CPLX>
__device__ __forceinline__ void tri_u(int j, int k, int p, float xr, float xi, float sr, float si,
                                      float &ur, float &ui) {
    const bool live = j > k + 1 && j < p;
    const float vr = CPLX ? xr * sr - xi * si : xr * sr;
    const float vi = CPLX ? xr * si + xi * sr : 0.f;
    ur = j == k + 1 ? 1.f : (live ? vr : 0.f);
    ui = live ? vi : 0.f;
}

// Sum over each group of G = 2 or 4 consecutive lanes, in every lane of the group.
template <int G>
__device__ __forceinline__ float group_sum(float v) {
    v += dppf<0xb1>(v);                         // quad_perm [1, 0, 3, 2]
    if constexpr (G == 4) v += dppf<0x4e>(v);   // quad_perm [2, 3, 0, 1]
    return v;
}

// T = Q^H A Q for the p x p Hermitian (or real symmetric) A, column-major with both triangles,
// by Householder reflections (LAPACK ssytd2 / chetd2, lower form): reflection k,
// H_k = I - tau_k u_k u_k^H with u_k[k + 1] = 1 and u_k[i] = 0 for i <= k, turns column k
// below the diagonal into (beta_k, 0, ..., 0) with beta_k real, and Q = H_0 H_1 ... H_{m-1}
// (m = p - 2 real; m = p - 1 complex, the last reflection making A[p-1][p-2] real).  T is real
// symmetric either way.
// Layout: the G = 4 lanes G i .. G i + G - 1 hold row i of A (P = G NC rows), lane g of them
// the NC columns g NC .. g NC + NC - 1 in registers, so the product A u needs only a sum over
// those lanes.
// Every wave keeps its own copy of u and w in LDS (written and read by the wave alone, so no
// barrier), zero outside k+1 .. p-1, which keeps the product and the rank-2 update free of
// masks.  Column k below the diagonal is row k conjugated (A stays Hermitian): the lanes of row
// k + 1 store it after the update, double-buffered by step parity.  Two barriers a step: before
// the reflector (every wave forms the same one from the stored row with its own DPP
// reductions) and after the product.  The reflectors, D, E and tau collect in LDS (dynamic:
// the packed reflectors take p (p - 1) / 2 elements) and go to memory once at the end.
// A is first scaled by the power of two that brings its largest magnitude into [0.5, 1)
// (exact, and eigenvectors do not change), so squares neither overflow nor underflow for Gram
// matrices of data far from unit magnitude.
// Out: D (p) and E (p - 1) of the scaled T, tau (m; complex as pairs) and the reflectors packed
// by column (u_k[k + 1 .. p - 1] from element k p - k (k + 1) / 2 of Upk).  stat[0..7] (status
// words of the kernels that follow) are cleared.
template <int NC, int G, bool CPLX>
__global__ __launch_bounds__(G * G * NC) void rank_tridiag_kernel(const float *__restrict__ Ain,
                                                               int p, float *__restrict__ D,
                                                               float *__restrict__ E,
                                                               float *__restrict__ Upk,
                                                               float *__restrict__ tau,
                                                               int *__restrict__ stat) {
    constexpr int P = G * NC, NT = G * P, NW = NT / 64, C = CPLX ? 2 : 1;
    constexpr int NS = (P + 63) / 64;   // row slots per lane in the wave-wide loops
    constexpr int L = 64 * NS;
    __shared__ float xs[2][C][L], ps[C][L];
    __shared__ __align__(16) float uw[NW][2][C][L];   // each wave's u and w
    __shared__ float red[NW];
    __shared__ float Ds[P], Es[P], Ts[C * P];
    extern __shared__ float Us[];   // p (p - 1) / 2 packed reflector elements
    const int tid = threadIdx.x, row = tid / G, j0 = (tid % G) * NC, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (tid < 8) stat[tid] = 0;
    for (int i = tid; i < L; i += NT) {
#pragma unroll
        for (int c = 0; c < C; ++c) xs[0][c][i] = xs[1][c][i] = ps[c][i] = 0.f;
    }
    float ar[NC], ai[CPLX ? NC : 1];
    float amax = 0.f;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int j = j0 + c;
        const bool in = row < p && j < p;
        const size_t e = (size_t)j * p + row;
        ar[c] = in ? Ain[C * e] : 0.f;
        amax = fmaxf(amax, fabsf(ar[c]));
        if constexpr (CPLX) {
            ai[c] = in ? Ain[2 * e + 1] : 0.f;
            amax = fmaxf(amax, fabsf(ai[c]));
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o));
    if (lane == 0) red[wave] = amax;
    __syncthreads();
    amax = red[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) amax = fmaxf(amax, red[w]);
    if (amax > 0.f && amax <= 3.4e38f) {
        int e;
        (void)frexpf(amax, &e);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            ar[c] = ldexpf(ar[c], -e);
            if constexpr (CPLX) ai[c] = ldexpf(ai[c], -e);
        }
    }
    // row r, conjugated (column r), into xs[b] by the lanes that hold it
    const auto put_row = [&](int r, int b) {
        if (row != r) return;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            xs[b][0][j0 + c] = ar[c];
            if constexpr (CPLX) xs[b][1][j0 + c] = -ai[c];
        }
    };
    float *ur = uw[wave][0][0], *ui = uw[wave][0][C - 1];
    float *wr = uw[wave][1][0], *wi = uw[wave][1][C - 1];
    put_row(0, 0);
    const int steps = CPLX ? p - 1 : p - 2;
    for (int k = 0; k < steps; ++k) {
        __syncthreads();
        const float *xr = xs[k & 1][0], *xi = xs[k & 1][C - 1];
        // u = (x - beta e_1) / (alpha - beta) for x = A[k+1:, k], alpha = x[0],
        // beta = -sign(Re alpha) |x|, tau = (beta - Re alpha) / beta - i Im alpha / beta; no
        // reflection (tau = 0) when x is already a real multiple of e_1.  Every wave reduces
        // the same values in the same order.
        const float alr = xr[k + 1], ali = CPLX ? xi[k + 1] : 0.f;
        float ss = 0.f;
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) {
            const int i = lane + 64 * sl;
            const bool live = i > k + 1 && i < p;
            const float vr = xr[i];
            ss = fmaf(vr, live ? vr : 0.f, ss);
            if constexpr (CPLX) {
                const float vi = xi[i];
                ss = fmaf(vi, live ? vi : 0.f, ss);
            }
        }
        ss = wave_total(ss);
        float tr = 0.f, tim = 0.f, beta = alr, sr = 0.f, si = 0.f;
        if (ss != 0.f || ali != 0.f) {
            const float nrm = sqrtf(fmaf(alr, alr, fmaf(ali, ali, ss)));
            const float b = alr >= 0.f ? -nrm : nrm;
            const float rb = 1.f / b;
            tr = (b - alr) * rb;
            tim = -ali * rb;
            const float dr = alr - b;   // 1 / (alpha - beta)
            const float rd = 1.f / fmaf(dr, dr, ali * ali);
            sr = dr * rd;
            si = -ali * rd;
            beta = b;
        }
        // this wave's u
#pragma unroll
        for (int sl = 0; sl < NS; ++sl) {
            const int i = lane + 64 * sl;
            float vr, vi;
            tri_u<CPLX>(i, k, p, xr[i], CPLX ? xi[i] : 0.f, sr, si, vr, vi);
            ur[i] = vr;
            if constexpr (CPLX) ui[i] = vi;
        }
        if (wave == 0) {
            const int o = packed_column(p, k) - (k + 1);
            for (int i = k + 1 + lane; i < p; i += 64) {
                Us[C * (o + i)] = ur[i];
                if constexpr (CPLX) Us[2 * (o + i) + 1] = ui[i];
            }
            if (lane == 0) {
                Ds[k] = xr[k];
                Es[k] = beta;
                Ts[C * k] = tr;
                if constexpr (CPLX) Ts[2 * k + 1] = tim;
            }
        }
        if (tr != 0.f || tim != 0.f) {   // the same values in every wave
            // ps = tau A' u: the quad of row i sums its four column ranges (16-byte reads of
            // this wave's u: j0 and the arrays are 32-byte aligned)
            float s_r = 0.f, s_i = 0.f;
            const float4 *ur4 = reinterpret_cast<const float4 *>(ur + j0);
            const float4 *ui4 = reinterpret_cast<const float4 *>(ui + j0);
#pragma unroll
            for (int c4 = 0; c4 < NC / 4; ++c4) {
                const float4 vr4 = ur4[c4];
                const float ucr[4] = {vr4.x, vr4.y, vr4.z, vr4.w};
                float uci[4] = {0.f, 0.f, 0.f, 0.f};
                if constexpr (CPLX) {
                    const float4 vi4 = ui4[c4];
                    uci[0] = vi4.x;
                    uci[1] = vi4.y;
                    uci[2] = vi4.z;
                    uci[3] = vi4.w;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int c = 4 * c4 + q;
                    s_r = fmaf(ar[c], ucr[q], s_r);
                    if constexpr (CPLX) {
                        s_r = fmaf(-ai[c], uci[q], s_r);
                        s_i = fmaf(ar[c], uci[q], fmaf(ai[c], ucr[q], s_i));
                    }
                }
            }
            s_r = group_sum<G>(s_r);
            if constexpr (CPLX) s_i = group_sum<G>(s_i);
            if (j0 == 0) {
                if constexpr (CPLX) {
                    ps[0][row] = tr * s_r - tim * s_i;
                    ps[1][row] = tr * s_i + tim * s_r;
                } else {
                    ps[0][row] = tr * s_r;
                }
            }
            __syncthreads();
            // K = -(tau / 2)(ps^H u), then this wave's w = ps + K u (0 outside k+1 .. p-1)
            float dr = 0.f, di = 0.f;
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                const int i = lane + 64 * sl;
                const bool live = i > k && i < p;
                const float p_r = live ? ps[0][i] : 0.f;
                const float u_r = ur[i];
                dr = fmaf(p_r, u_r, dr);
                if constexpr (CPLX) {
                    const float p_i = live ? ps[1][i] : 0.f;
                    const float u_i = ui[i];
                    dr = fmaf(p_i, u_i, dr);
                    di = fmaf(p_r, u_i, fmaf(-p_i, u_r, di));
                }
            }
            dr = wave_total(dr);
            if constexpr (CPLX) di = wave_total(di);
            const float Kr = -0.5f * (tr * dr - tim * di);
            const float Ki = -0.5f * (tr * di + tim * dr);
#pragma unroll
            for (int sl = 0; sl < NS; ++sl) {
                const int i = lane + 64 * sl;
                const bool live = i > k && i < p;
                const float u_r = ur[i];
                if constexpr (CPLX) {
                    const float u_i = ui[i];
                    wr[i] = live ? ps[0][i] + Kr * u_r - Ki * u_i : 0.f;
                    wi[i] = live ? ps[1][i] + Kr * u_i + Ki * u_r : 0.f;
                } else {
                    wr[i] = live ? ps[0][i] + Kr * u_r : 0.f;
                }
            }
            // row i: A[i][j] -= u_i conj(w_j) + w_i conj(u_j)
            const float uir = ur[row], wir = wr[row];
            const float uii = CPLX ? ui[row] : 0.f, wii = CPLX ? wi[row] : 0.f;
            const float4 *wr4 = reinterpret_cast<const float4 *>(wr + j0);
            const float4 *wi4 = reinterpret_cast<const float4 *>(wi + j0);
#pragma unroll
            for (int c4 = 0; c4 < NC / 4; ++c4) {
                const float4 a4 = ur4[c4], b4 = wr4[c4];
                const float ujr[4] = {a4.x, a4.y, a4.z, a4.w}, wjr[4] = {b4.x, b4.y, b4.z, b4.w};
                float uji[4] = {0.f, 0.f, 0.f, 0.f}, wji[4] = {0.f, 0.f, 0.f, 0.f};
                if constexpr (CPLX) {
                    const float4 c4i = ui4[c4], d4 = wi4[c4];
                    uji[0] = c4i.x;
                    uji[1] = c4i.y;
                    uji[2] = c4i.z;
                    uji[3] = c4i.w;
                    wji[0] = d4.x;
                    wji[1] = d4.y;
                    wji[2] = d4.z;
                    wji[3] = d4.w;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int c = 4 * c4 + q;
                    float r = fmaf(uir, wjr[q], wir * ujr[q]);
                    if constexpr (CPLX) {
                        r = fmaf(uii, wji[q], fmaf(wii, uji[q], r));
                        const float m = fmaf(uii, wjr[q],
                                             fmaf(-uir, wji[q], fmaf(wii, ujr[q], -wir * uji[q])));
                        ai[c] -= m;
                    }
                    ar[c] -= r;
                }
            }
        }
        put_row(k + 1, (k + 1) & 1);
    }
    // the last rows: D[p-1] (and D[p-2], E[p-2] when real, whose last step is empty)
    __syncthreads();
    if (!CPLX) put_row(p - 1, (p - 1) & 1);
    __syncthreads();
    if (tid == 0) {
        Ds[p - 1] = xs[(p - 1) & 1][0][p - 1];
        if (!CPLX) {
            Ds[p - 2] = xs[(p - 2) & 1][0][p - 2];
            Es[p - 2] = xs[(p - 2) & 1][0][p - 1];
        }
    }
    __syncthreads();
    const int npk = C * (p * (p - 1) / 2);
    for (int e = tid; e < npk; e += NT) Upk[e] = Us[e];
    for (int i = tid; i < p; i += NT) {
        D[i] = Ds[i];
        if (i < p - 1) E[i] = Es[i];
    }
    for (int i = tid; i < C * steps; i += NT) tau[i] = Ts[i];
}

// Number of eigenvalues of the symmetric tridiagonal T below x, by the signs of the pivots of
// T - x I (a pivot of magnitude <= pivmin is taken as -pivmin, LAPACK's guard); de[i] holds
// (d[i], e[i-1]^2) with e[-1] = 0, and (1e30, 0) past p (pivots that add nothing to the count),
// read eight entries ahead of the recurrence.
__device__ __forceinline__ int sturm_count(const float2 *de, int p, float x, float pivmin) {
    float q = 1.f;
    int c = 0;
    for (int i0 = 0; i0 < p; i0 += 8) {
        float2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = de[i0 + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            q = (v[u].x - x) - __fdividef(v[u].y, q);
            q = fabsf(q) <= pivmin ? -pivmin : q;
            c += q < 0.f;
        }
    }
    return c;
}

// A fixed pseudo-random start vector entry in (-1, 1) for row i of eigenvector index j.
__device__ __forceinline__ float start_entry(int j, int i) {
    uint32_t h = (uint32_t)j * 0x9E3779B1u ^ ((uint32_t)i + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return (float)(int32_t)h * (1.f / 2147483648.f);
}

constexpr int kEigWaves = 8;

// Eigenpairs of the symmetric tridiagonal T (diagonal D, off-diagonal E, p <= PMAX) for the r
// eigenvalue indices the projection reads: ascending index c_first + q for q < q_split,
// c_second + (q - q_split) after.  Eigenvalues by multisection of Sturm counts (each lane of
// a wave counts at one of 64 points of the bracket, narrowing it 65-fold a round, to within
// 4 eps |T|); eigenvectors by inverse iteration as LAPACK's sstein does it: an LU of
// T - lambda I with partial pivoting (pivots below eps |T| raised to it), three solves from a
// fixed pseudo-random start, and after each solve Gram-Schmidt against the earlier vectors of
// the same cluster (consecutive eigenvalues closer than 1e-3 |T|), whose eigenvalues are
// first moved apart by at least 10 eps |lambda|.  One workgroup of 8 waves: wave w brackets
// eigenvalues w, w + 8, ..., then runs clusters w, w + 8, ... (the LU and the solves on its
// lane 0, eight entries of LDS read ahead of the recurrences; products and norms across the
// wave).  Y (p x r): column q is the unit eigenvector of index idx(q); T = 0 gives unit
// vectors.  A non-finite vector sets stat[0].
template <int PMAX>
__global__ __launch_bounds__(64 * kEigWaves) void rank_tridiag_eigvec_kernel(
    const float *__restrict__ D, const float *__restrict__ E, int p, int r, int q_split,
    int c_first, int c_second, float *__restrict__ Y, int *__restrict__ stat) {
    constexpr int PP = PMAX + 8;   // room for the read-ahead past p
    __shared__ float2 de[PP];      // (d[i], e[i-1]^2)
    __shared__ float sd[PP], se[PP], lam[PMAX];
    __shared__ int cstart[PMAX + 1];
    __shared__ float4 lu[kEigWaves][PP];   // (1 / U diagonal, U's two superdiagonals, L)
    __shared__ unsigned char piv[kEigWaves][PP];   // rows k and k + 1 interchanged
    __shared__ float xv[kEigWaves][PP];
    __shared__ float red[kEigWaves][4];
    __shared__ int ncl;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr float eps = 1.1920929e-7f;
    for (int i = tid; i < PP; i += 64 * kEigWaves) {
        const float d = i < p ? D[i] : 0.f;
        const float e = i < p - 1 ? E[i] : 0.f;
        const float em = i > 0 && i < p ? E[i - 1] : 0.f;
        sd[i] = d;
        se[i] = e;
        de[i] = i < p ? make_float2(d, em * em) : make_float2(1e30f, 0.f);
    }
    __syncthreads();
    // |T| (row sums), Gershgorin bounds and the pivot guard
    float nrm = 0.f, lo = 3.4e38f, hi = -3.4e38f, e2max = 0.f;
    for (int i = tid; i < p; i += 64 * kEigWaves) {
        const float off = fabsf(se[i]) + (i > 0 ? fabsf(se[i - 1]) : 0.f);
        nrm = fmaxf(nrm, fabsf(sd[i]) + off);
        lo = fminf(lo, sd[i] - off);
        hi = fmaxf(hi, sd[i] + off);
        e2max = fmaxf(e2max, se[i] * se[i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        nrm = fmaxf(nrm, __shfl_xor(nrm, o));
        lo = fminf(lo, __shfl_xor(lo, o));
        hi = fmaxf(hi, __shfl_xor(hi, o));
        e2max = fmaxf(e2max, __shfl_xor(e2max, o));
    }
    if (lane == 0) {
        red[wave][0] = nrm;
        red[wave][1] = lo;
        red[wave][2] = hi;
        red[wave][3] = e2max;
    }
    __syncthreads();
    nrm = red[0][0];
    lo = red[0][1];
    hi = red[0][2];
    e2max = red[0][3];
#pragma unroll
    for (int w = 1; w < kEigWaves; ++w) {
        nrm = fmaxf(nrm, red[w][0]);
        lo = fminf(lo, red[w][1]);
        hi = fmaxf(hi, red[w][2]);
        e2max = fmaxf(e2max, red[w][3]);
    }
    const auto index_of = [&](int q) {
        return q < q_split ? c_first + q : c_second + (q - q_split);
    };
    if (nrm == 0.f) {   // T = 0: every vector is an eigenvector
        for (int e = tid; e < p * r; e += 64 * kEigWaves) {
            const int q = e / p, i = e - q * p;
            Y[e] = i == index_of(q) ? 1.f : 0.f;
        }
        return;
    }
    const float pivmin = 1.17549435e-38f * fmaxf(1.f, e2max);
    const float widen = 4.f * eps * (float)p * nrm + 2.f * pivmin;
    lo -= widen;
    hi += widen;
    const float tol = 4.f * eps * nrm;
    // eigenvalue idx: count(a) <= idx < count(b) throughout
    for (int q = wave; q < r; q += kEigWaves) {
        const int idx = index_of(q);
        float a = lo, b = hi;
        for (int round = 0; round < 8 && b - a > tol; ++round) {
            const float x = a + (b - a) * (float)(lane + 1) * (1.f / 65.f);
            const int c = sturm_count(de, p, x, pivmin);
            const uint64_t above = __ballot(c > idx);
            const int f = above ? __builtin_ctzll(above) : 64;
            const float xf = __shfl(x, f < 64 ? f : 63);
            const float xp = __shfl(x, f > 0 ? f - 1 : 0);
            const float na = f > 0 ? xp : a, nb = f < 64 ? xf : b;
            a = na;
            b = nb;
        }
        if (lane == 0) lam[q] = 0.5f * (a + b);
    }
    __syncthreads();
    // clusters of consecutive requested eigenvalues (ascending in q)
    const float ortol = 1e-3f * nrm;
    if (tid == 0) {
        int n = 0;
        for (int q = 0; q < r; ++q)
            if (q == 0 || lam[q] - lam[q - 1] > ortol) cstart[n++] = q;
        cstart[n] = r;
        ncl = n;
    }
    __syncthreads();
    const float tolpiv = eps * nrm;
    float4 *f = lu[wave];
    unsigned char *pv = piv[wave];
    float *x = xv[wave];
    bool bad = false;
    for (int cl = wave; cl < ncl; cl += kEigWaves) {
        const int q0 = cstart[cl], q1 = cstart[cl + 1];
        float prev = 0.f;
        for (int q = q0; q < q1; ++q) {
            float lm = lam[q];
            if (q > q0) lm = fmaxf(lm, prev + 10.f * eps * fabsf(prev));
            prev = lm;
            const int kb = (p + 7) & ~7;   // the loops below run whole blocks of 8
            if (lane == 0) {
                // LU of T - lm I with partial pivoting (LAPACK slagtf), branch-free; entries
                // from p - 1 on are (0, 0, 0, 0) without interchange (inert in the solves),
                // then the last pivot goes to p - 1
                const auto guard = [&](float v) {
                    return fabsf(v) < tolpiv ? (v < 0.f ? -tolpiv : tolpiv) : v;
                };
                float a = sd[0] - lm, b = se[0], afin = a;
                for (int k0 = 0; k0 < kb; k0 += 8) {
                    float dn[8], en[9];
#pragma unroll
                    for (int u = 0; u < 8; ++u) dn[u] = sd[k0 + u + 1];
#pragma unroll
                    for (int u = 0; u < 9; ++u) en[u] = se[k0 + u];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int k = k0 + u;
                        const float c = en[u];
                        const float an = dn[u] - lm;
                        const float bn = en[u + 1];   // se is 0 from p - 1 on
                        const bool swap = fabsf(a) < fabsf(c);
                        const float m1 = a != 0.f ? c * __builtin_amdgcn_rcpf(a) : 0.f;
                        const float m2 = a * __builtin_amdgcn_rcpf(c);
                        const bool live = k < p - 1;
                        const float m = swap ? m2 : m1;
                        float4 o = swap ? make_float4(c, an, bn, m) : make_float4(a, b, 0.f, m);
                        o.x = __builtin_amdgcn_rcpf(guard(o.x));
                        f[k] = live ? o : make_float4(0.f, 0.f, 0.f, 0.f);
                        pv[k] = live && swap;
                        const float na = swap ? b - m * an : an - m * b;
                        b = swap ? -m * bn : bn;
                        a = na;
                        afin = k == p - 2 ? a : afin;
                    }
                }
                f[p - 1] = make_float4(__builtin_amdgcn_rcpf(guard(afin)), 0.f, 0.f, 0.f);
            }
            for (int i = lane; i < PP; i += 64)
                x[i] = i < p ? start_entry(index_of(q), i) : 0.f;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            for (int it = 0; it < 3; ++it) {
                if (lane == 0) {
                    // forward: the row operations of the factorisation (x is 0 from p on)
                    float cur = x[0];
                    for (int k0 = 0; k0 < kb; k0 += 8) {
                        float xn[8], mn[8];
                        bool sw[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            xn[u] = x[k0 + u + 1];
                            mn[u] = f[k0 + u].w;
                            sw[u] = pv[k0 + u] != 0;
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const float a0 = sw[u] ? xn[u] : cur;
                            const float a1 = sw[u] ? cur : xn[u];
                            x[k0 + u] = a0;
                            cur = a1 - mn[u] * a0;
                        }
                    }
                    // back substitution from the padded top (the first component of f holds
                    // the reciprocal pivot; f and x are 0 past p - 1)
                    float x1 = 0.f, x2 = 0.f;
                    for (int k1 = kb - 1; k1 >= 0; k1 -= 8) {
                        float xk[8];
                        float4 fk[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            xk[u] = x[k1 - u];
                            fk[u] = f[k1 - u];
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const float v = (xk[u] - fk[u].y * x1 - fk[u].z * x2) * fk[u].x;
                            x[k1 - u] = v;
                            x2 = x1;
                            x1 = v;
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                // scale by the largest magnitude first (no overflow in the products below)
                float mx = 0.f;
                for (int i = lane; i < p; i += 64) mx = fmaxf(mx, fabsf(x[i]));
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
                const float inv = mx > 0.f && mx <= 3.4e38f ? 1.f / mx : 0.f;
                float xl[PMAX / 64];
#pragma unroll
                for (int e = 0; e < PMAX / 64; ++e) {
                    const int i = lane + 64 * e;
                    xl[e] = i < p ? x[i] * inv : 0.f;
                }
                // Gram-Schmidt against the earlier vectors of the cluster (unit, in Y)
                for (int qq = q0; qq < q; ++qq) {
                    const float *yq = Y + (size_t)qq * p;
                    float yv[PMAX / 64];
                    float s = 0.f;
#pragma unroll
                    for (int e = 0; e < PMAX / 64; ++e) {
                        const int i = lane + 64 * e;
                        yv[e] = i < p ? yq[i] : 0.f;
                        s = fmaf(xl[e], yv[e], s);
                    }
                    s = wave_total(s);
#pragma unroll
                    for (int e = 0; e < PMAX / 64; ++e) xl[e] = fmaf(-s, yv[e], xl[e]);
                }
                float ss = 0.f;
#pragma unroll
                for (int e = 0; e < PMAX / 64; ++e) ss = fmaf(xl[e], xl[e], ss);
                ss = wave_total(ss);
                const float sc = ss > 0.f ? 1.f / sqrtf(ss) : 0.f;
                if (!(ss > 0.f) || !(ss <= 3.4e38f)) bad = true;
#pragma unroll
                for (int e = 0; e < PMAX / 64; ++e) {
                    const int i = lane + 64 * e;
                    if (i < p) x[i] = xl[e] * sc;
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
            for (int i = lane; i < p; i += 64) Y[(size_t)q * p + i] = x[i];
        }
    }
    if (bad && lane == 0) atomicOr(stat, 1);
}

// Eigenvectors of A from those of its tridiagonal form: z = H_0 H_1 ... H_{m-1} y, the
// reflections applied last to first (z -= tau_k u_k (u_k^H z)), one wave per vector with NE
// elements per lane (p <= 64 NE).  The workgroup first copies the packed reflectors into LDS
// (dynamic, p (p - 1) / 2 elements), then each wave reads the next reflector while it applies
// the current one.  Vector q (column q of the real Y) goes to column c_first + q, or
// c_second + (q - q_split) for q >= q_split, of V (column-major p x p, complex as pairs),
// where the projection reads it.  *info = stat[0] (set by the eigenpair kernel when a vector
// is not finite).
template <int NE, bool CPLX>
__global__ __launch_bounds__(256) void rank_backtransform_kernel(
    const float *__restrict__ Upk, const float *__restrict__ tau, const float *__restrict__ Y,
    int p, int r, int q_split, int c_first, int c_second, const int *__restrict__ stat,
    float *__restrict__ V, int *__restrict__ info) {
    constexpr int C = CPLX ? 2 : 1;
    extern __shared__ float Us[];
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (blockIdx.x == 0 && threadIdx.x == 0) *info = stat[0] != 0 ? 1 : 0;
    const int npk = C * (p * (p - 1) / 2);
    for (int e = threadIdx.x; e < npk; e += 256) Us[e] = Upk[e];
    __syncthreads();
    if (q >= r) return;
    const float *y = Y + (size_t)q * p;
    float zr[NE], zi[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const int i = lane + 64 * e;
        zr[e] = i < p ? y[i] : 0.f;
        zi[e] = 0.f;
    }
    // u_k[i] for i > k (0 below), loaded one reflector ahead
    const auto load = [&](int k, float (&ur)[NE], float (&ui)[NE]) {
        const int o = packed_column(p, k) - (k + 1);
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int i = lane + 64 * e;
            const bool in = k >= 0 && i > k && i < p;
            ur[e] = in ? Us[C * (o + i)] : 0.f;
            ui[e] = in && CPLX ? Us[C * (o + i) + C - 1] : 0.f;
        }
    };
    const int kt = CPLX ? p - 2 : p - 3;   // the last reflection
    float unr[NE], uni[NE];
    float tnr = kt >= 0 ? tau[C * kt] : 0.f, tni = kt >= 0 && CPLX ? tau[C * kt + C - 1] : 0.f;
    load(kt, unr, uni);
    for (int k = kt; k >= 0; --k) {
        float ur[NE], ui[NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            ur[e] = unr[e];
            ui[e] = uni[e];
        }
        const float t_r = tnr, t_i = tni;
        tnr = k >= 1 ? tau[C * (k - 1)] : 0.f;
        tni = k >= 1 && CPLX ? tau[C * (k - 1) + C - 1] : 0.f;
        load(k - 1, unr, uni);
        // s = u^H z
        float sre = 0.f, sim = 0.f;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            sre = fmaf(ur[e], zr[e], sre);
            if constexpr (CPLX) {
                sre = fmaf(ui[e], zi[e], sre);
                sim = fmaf(ur[e], zi[e], fmaf(-ui[e], zr[e], sim));
            }
        }
        sre = wave_total(sre);
        if constexpr (CPLX) sim = wave_total(sim);
        // z -= (tau s) u
        const float ar = t_r * sre - t_i * sim, ai = t_r * sim + t_i * sre;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            zr[e] = fmaf(-ar, ur[e], zr[e]);
            if constexpr (CPLX) {
                zr[e] = fmaf(ai, ui[e], zr[e]);
                zi[e] = fmaf(-ar, ui[e], fmaf(-ai, ur[e], zi[e]));
            }
        }
    }
    const int col = q < q_split ? c_first + q : c_second + (q - q_split);
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const int i = lane + 64 * e;
        if (i < p) {
            const size_t o = (size_t)col * p + i;
            V[C * o] = zr[e];
            if constexpr (CPLX) V[2 * o + 1] = zi[e];
        }
    }
}

}  // namespace

void rank_plan_launch_shape(RankPlan &pl, int cu_count) {
    pl.pp = (pl.p + 15) / 16 * 16;
    // FP32 with contiguous aligned Gram vectors and p <= 256: the LDS-staged kernel
    // (FFDAS_RANK_GRAM=wave selects the per-wave block kernel instead)
    const char *gk = std::getenv("FFDAS_RANK_GRAM");
    // (complex up to p = 208: its 17-tile variant would spill registers)
    pl.lds_gram = !pl.dbl && pl.gvec && pl.p > 16 && pl.pp <= (pl.cplx ? 208 : 256) &&
                  !(gk && std::string(gk) == "wave");
    // it needs at least one workgroup per CU with 256 k each (measured: at 200 x 36864
    // complex the block kernel is faster, at 128 x 131072 real the LDS kernel)
    if (pl.lds_gram && pl.L < 256LL * std::max(cu_count, 1)) pl.lds_gram = false;
    if (pl.lds_gram) {
        const int nt = pl.pp / 16;
        const int64_t ntri = (int64_t)nt * (nt + 1) / 2;
        const int64_t esz = pl.cplx ? 8 : 4;
        // two workgroups of 8 waves per CU, whole 32-k stages, at least 256 k each
        int64_t splits = std::max<int64_t>(1, 2LL * std::max(cu_count, 1));
        splits = std::min<int64_t>(splits, std::max<int64_t>(1, pl.L / 256));
        int64_t kchunk = (pl.L + splits - 1) / splits;
        kchunk = (kchunk + 31) / 32 * 32;
        pl.kchunk = kchunk;
        pl.splits = (int)((pl.L + kchunk - 1) / kchunk);
        pl.bs = 16;
        pl.nb = nt;
        pl.part_bytes = (size_t)(pl.splits * ntri * 256 * esz);
    }
    // measured on gfx950 (64 and 200 frames, real and complex FP32): single-tile blocks with
    // many waves for p <= 64, 32 x 32 blocks above; 64 x 64 blocks were slower everywhere
    if (!pl.lds_gram) {
    pl.bs = pl.p <= 64 ? 16 : 32;
    if (const char *e = std::getenv("FFDAS_RANK_BS")) {
        const int v = std::atoi(e);
        if (v == 16 || v == 32 || v == 64) pl.bs = v;
    }
    int waves_per_cu = 16;
    if (const char *e = std::getenv("FFDAS_RANK_WAVES")) waves_per_cu = std::max(1, std::atoi(e));
    pl.nb = (pl.p + pl.bs - 1) / pl.bs;
    const int64_t nblk = (int64_t)pl.nb * (pl.nb + 1) / 2;
    const int64_t esz = (pl.dbl ? 8 : 4) * (pl.cplx ? 2 : 1);
    // about eight waves per CU, at least 256 k per wave, partials within 256 MiB
    int64_t splits = std::max<int64_t>(
        1, ((int64_t)waves_per_cu * std::max(cu_count, 1) + nblk - 1) / nblk);
    splits = std::min<int64_t>(splits, std::max<int64_t>(1, pl.L / 256));
    const int64_t cap = (256LL << 20) / (nblk * pl.bs * pl.bs * esz);
    splits = std::max<int64_t>(1, std::min<int64_t>(splits, cap));
    int64_t kchunk = (pl.L + splits - 1) / splits;
    kchunk = (kchunk + 15) / 16 * 16;
    pl.kchunk = kchunk;
    // a multiple of 8 (one k-range per XCD in turn); trailing k-ranges may be empty
    pl.splits = (int)(((pl.L + kchunk - 1) / kchunk + 7) / 8 * 8);
    pl.part_bytes = (size_t)(pl.splits * nblk * pl.bs * pl.bs * esz);
    }
    const int tiles = (pl.r + 15) / 16;
    if (tiles <= 8) {
        pl.rt = tiles <= 1 ? 1 : (tiles <= 2 ? 2 : (tiles <= 4 ? 4 : 8));
        pl.passes = 1;
    } else {
        pl.rt = 8;
        pl.passes = (tiles + 7) / 8;
    }
    pl.rp = 16 * pl.rt * pl.passes;
}

ffdas_error_t launch_rank_gram(hipStream_t st, const RankPlan &pl, const void *x, void *part,
                               void *H) {
    if (pl.dbl && pl.cplx)
        gram_dispatch<double, true>(st, pl, x, part, H);
    else if (pl.dbl)
        gram_dispatch<double, false>(st, pl, x, part, H);
    else if (pl.cplx)
        gram_dispatch<float, true>(st, pl, x, part, H);
    else
        gram_dispatch<float, false>(st, pl, x, part, H);
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

bool rank_jacobi_fits(const RankPlan &pl, int lds_bytes) {
    const int pe = (pl.p + 1) / 2 * 2;
    const size_t esz = (pl.dbl ? 8 : 4) * (pl.cplx ? 2 : 1);
    // the kernel's static LDS (reduction slots, flags) stays well under 512 bytes; its V
    // registers hold ceil(pe / row slots) rows per thread
    const int slots = pe <= 64 ? 32 : 16;
    return pl.p >= 2 && (pe + slots - 1) / slots <= jacobi_vrows(pe <= 64 ? 32 : 64, esz) &&
           jacobi_lds_bytes(pe, esz, pl.dbl ? 8 : 4) + 512 <= (size_t)lds_bytes;
}

ffdas_error_t launch_rank_jacobi(hipStream_t st, const RankPlan &pl, void *H, int *info,
                                 int max_sweeps) {
    const int pe = (pl.p + 1) / 2 * 2;
    const size_t esz = (pl.dbl ? 8 : 4) * (pl.cplx ? 2 : 1);
    const size_t lds = jacobi_lds_bytes(pe, esz, pl.dbl ? 8 : 4);
    ProfScope ps(st, "rank_jacobi");
#define FFDAS_RANK_JACOBI(T, C)                                                               \
    do {                                                                                     \
        if (pe <= 64) {                                                                      \
            (void)hipFuncSetAttribute((const void *)rank_jacobi_kernel<T, C, 32>,            \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
            hipLaunchKernelGGL((rank_jacobi_kernel<T, C, 32>), dim3(1), dim3(1024), lds, st,   \
                               (T *)H, pl.p, pe, max_sweeps, info);                          \
        } else {                                                                             \
            (void)hipFuncSetAttribute((const void *)rank_jacobi_kernel<T, C, 64>,            \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
            hipLaunchKernelGGL((rank_jacobi_kernel<T, C, 64>), dim3(1), dim3(1024), lds, st,   \
                               (T *)H, pl.p, pe, max_sweeps, info);                          \
        }                                                                                    \
    } while (0)
    if (pl.dbl && pl.cplx)
        FFDAS_RANK_JACOBI(double, true);
    else if (pl.dbl)
        FFDAS_RANK_JACOBI(double, false);
    else if (pl.cplx)
        FFDAS_RANK_JACOBI(float, true);
    else
        FFDAS_RANK_JACOBI(float, false);
#undef FFDAS_RANK_JACOBI
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

ffdas_error_t launch_rank_project(hipStream_t st, const RankPlan &pl, const void *V,
                                  const void *x, void *out, void *Qh, void *Qn) {
    const int64_t n = (int64_t)pl.rp * pl.pp;
    const unsigned blocks = (unsigned)((n + 255) / 256);
#define FFDAS_RANK_SELECT(T, C)                                                              \
    hipLaunchKernelGGL((rank_select_kernel<T, C>), dim3(blocks), dim3(256), 0, st,          \
                       (const T *)V, pl.p, pl.pp, pl.r, pl.rp, pl.q_split, pl.c_first,       \
                       pl.c_second, (T *)Qh, (T *)Qn)
    if (pl.dbl) {
        if (pl.cplx) {
            FFDAS_RANK_SELECT(double, true);
            project_dispatch<double, true>(st, pl, x, out, Qh, Qn);
        } else {
            FFDAS_RANK_SELECT(double, false);
            project_dispatch<double, false>(st, pl, x, out, Qh, Qn);
        }
    } else {
        if (pl.cplx) {
            FFDAS_RANK_SELECT(float, true);
            project_dispatch<float, true>(st, pl, x, out, Qh, Qn);
        } else {
            FFDAS_RANK_SELECT(float, false);
            project_dispatch<float, false>(st, pl, x, out, Qh, Qn);
        }
    }
#undef FFDAS_RANK_SELECT
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

ffdas_error_t launch_rank_copy(hipStream_t st, bool dbl, bool cplx, const void *x, int64_t xs0,
                               int64_t xs1, void *out, int64_t os0, int64_t os1, int64_t m,
                               int64_t n) {
    const int64_t tot = m * n;
    if (tot == 0) return FFDAS_SUCCESS;
    const unsigned blocks = (unsigned)((tot + 255) / 256);
#define FFDAS_RANK_COPY(T, C)                                                                 \
    hipLaunchKernelGGL((rank_copy_kernel<T, C>), dim3(blocks), dim3(256), 0, st, (const T *)x, \
                       xs0, xs1, (T *)out, os0, os1, m, n)
    if (dbl && cplx)
        FFDAS_RANK_COPY(double, true);
    else if (dbl)
        FFDAS_RANK_COPY(double, false);
    else if (cplx)
        FFDAS_RANK_COPY(float, true);
    else
        FFDAS_RANK_COPY(float, false);
#undef FFDAS_RANK_COPY
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}


size_t rank_tridiag_lds(int p, bool cplx) {
    return (size_t)p * (p - 1) / 2 * (cplx ? 8 : 4);
}

// Real p <= 224, complex p <= 160: a row of A fits the registers of its four lanes (56 real
// or 40 complex columns each); the tridiagonalisation's static LDS stays under 44 KiB.
int tridiag_max_p(bool cplx) { return cplx ? 160 : 224; }

bool rank_tridiag_fits(const RankPlan &pl, int lds_bytes) {
    return !pl.dbl && pl.p >= 2 && pl.p <= tridiag_max_p(pl.cplx) &&
           rank_tridiag_lds(pl.p, pl.cplx) + 45056 <= (size_t)lds_bytes;
}

ffdas_error_t launch_rank_tridiag(hipStream_t st, const RankPlan &pl, int lds_bytes,
                                  const void *H, void *D, void *E, void *U, void *tau,
                                  int *stat) {
    if (!rank_tridiag_fits(pl, lds_bytes)) return FFDAS_ERROR_INVALID_ARGUMENT;
    ProfScope ps(st, "rank_tridiag");
    const size_t lds = rank_tridiag_lds(pl.p, pl.cplx);
#define FFDAS_RANK_TRIDIAG(NC, G, CPLX)                                                        \
    do {                                                                                       \
        (void)hipFuncSetAttribute((const void *)rank_tridiag_kernel<NC, G, CPLX>,              \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);       \
        hipLaunchKernelGGL((rank_tridiag_kernel<NC, G, CPLX>), dim3(1), dim3(G * G * NC), lds, \
                           st, (const float *)H, pl.p, (float *)D, (float *)E, (float *)U,     \
                           (float *)tau, stat);                                                \
    } while (0)
    // P rows >= p, 4 lanes a row of NC (a multiple of 8) columns each.  Two lanes a row (half
    // the waves, so half the per-step work every wave repeats) measured slower at p = 200 real
    // (0.53 against 0.50 ms): the step is bound by its chain of dependent phases, not by issue.
    const int nc8 = (pl.p + 31) / 32;
    if (pl.cplx) {
        switch (nc8) {
            case 1: FFDAS_RANK_TRIDIAG(8, 4, true); break;
            case 2: FFDAS_RANK_TRIDIAG(16, 4, true); break;
            case 3: FFDAS_RANK_TRIDIAG(24, 4, true); break;
            case 4: FFDAS_RANK_TRIDIAG(32, 4, true); break;
            default: FFDAS_RANK_TRIDIAG(40, 4, true); break;
        }
    } else {
        switch (nc8) {
            case 1: FFDAS_RANK_TRIDIAG(8, 4, false); break;
            case 2: FFDAS_RANK_TRIDIAG(16, 4, false); break;
            case 3: FFDAS_RANK_TRIDIAG(24, 4, false); break;
            case 4: FFDAS_RANK_TRIDIAG(32, 4, false); break;
            case 5: FFDAS_RANK_TRIDIAG(40, 4, false); break;
            case 6: FFDAS_RANK_TRIDIAG(48, 4, false); break;
            default: FFDAS_RANK_TRIDIAG(56, 4, false); break;
        }
    }
#undef FFDAS_RANK_TRIDIAG
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

ffdas_error_t launch_rank_tridiag_eigvec(hipStream_t st, const RankPlan &pl, const void *D,
                                         const void *E, void *Y, int *stat) {
    if (pl.dbl || pl.p < 2 || pl.p > tridiag_max_p(pl.cplx) || pl.r < 1 || pl.r > pl.p)
        return FFDAS_ERROR_INVALID_ARGUMENT;
    ProfScope ps(st, "rank_tridiag_eigvec");
#define FFDAS_RANK_EIGVEC(PMAX)                                                                \
    hipLaunchKernelGGL((rank_tridiag_eigvec_kernel<PMAX>), dim3(1), dim3(64 * kEigWaves), 0,   \
                       st, (const float *)D, (const float *)E, pl.p, pl.r, pl.q_split,         \
                       pl.c_first, pl.c_second, (float *)Y, stat)
    switch ((pl.p + 63) / 64) {
        case 1: FFDAS_RANK_EIGVEC(64); break;
        case 2: FFDAS_RANK_EIGVEC(128); break;
        case 3: FFDAS_RANK_EIGVEC(192); break;
        default: FFDAS_RANK_EIGVEC(256); break;
    }
#undef FFDAS_RANK_EIGVEC
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

ffdas_error_t launch_rank_backtransform(hipStream_t st, const RankPlan &pl, int lds_bytes,
                                        const void *U, const void *tau, const void *Y,
                                        const int *stat, void *V, int *info) {
    if (!rank_tridiag_fits(pl, lds_bytes) || pl.r < 1 || pl.r > pl.p)
        return FFDAS_ERROR_INVALID_ARGUMENT;
    ProfScope ps(st, "rank_backtransform");
    const unsigned blocks = (unsigned)((pl.r + 3) / 4);
    const size_t lds = rank_tridiag_lds(pl.p, pl.cplx);
#define FFDAS_RANK_BACK(NE, CPLX)                                                              \
    do {                                                                                       \
        (void)hipFuncSetAttribute((const void *)rank_backtransform_kernel<NE, CPLX>,           \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);       \
        hipLaunchKernelGGL((rank_backtransform_kernel<NE, CPLX>), dim3(blocks), dim3(256), lds, \
                           st, (const float *)U, (const float *)tau, (const float *)Y, pl.p,   \
                           pl.r, pl.q_split, pl.c_first, pl.c_second, stat, (float *)V, info); \
    } while (0)
    const int ne = (pl.p + 63) / 64;
    if (pl.cplx) {
        switch (ne) {
            case 1: FFDAS_RANK_BACK(1, true); break;
            case 2: FFDAS_RANK_BACK(2, true); break;
            default: FFDAS_RANK_BACK(3, true); break;
        }
    } else {
        switch (ne) {
            case 1: FFDAS_RANK_BACK(1, false); break;
            case 2: FFDAS_RANK_BACK(2, false); break;
            case 3: FFDAS_RANK_BACK(3, false); break;
            default: FFDAS_RANK_BACK(4, false); break;
        }
    }
#undef FFDAS_RANK_BACK
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

}  // namespace ffdas
