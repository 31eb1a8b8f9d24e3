// Interpolation from a structured grid with arbitrary vertex positions to scattered points
// (ffdas_create_interpolation_plan / _preprocess / ffdas_interpolation / _destroy).
//
// Plan (once per grid): the grid's bounding box (per-block partial min/max, finished on the
// host), 128 bins per axis over it, every tetrahedron counted into each bin of its quantised
// bounding box (atomics), a device scan of the counts, and a fill of the per-bin lists.  The
// list order inside a bin is whatever the atomics give: the query takes the lowest passing
// tetrahedron id of the whole bin, which is the reference's answer (its lists are sorted by
// id and it returns the first that passes) without sorting anything.
// Query (once per point set): per point, its bin and the FP32 barycentric test of each
// candidate, computed without contraction in the reference's operation order; the result is
// four vertex ids and weights (linear) or one vertex id (nearest).
// Evaluation (per call, the per-frame part): x is first copied frame-inner (the frames of a
// chunk, 16 for 4-byte data and 8 otherwise, contiguous per vertex), so one thread per query
// point fetches each of its vertices' chunk with one vector load; its lookup is read once per
// chunk; out is packed.
#include <hip/hip_runtime.h>

#include <cstring>   // before rocPRIM, whose iterators use memset

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "internal.h"

struct ffdas_interpolation_plan {
    int device = 0;
    int64_t nx = 0, ny = 0, nz = 0;
    int mode = FFDAS_INTERP_LINEAR;
    float *verts = nullptr;            // 3 * nx * ny * nz
    float4 *verts4 = nullptr;          // the same, one 16-byte vector per vertex (location)
    float bmin[3] = {0, 0, 0}, scale[3] = {0, 0, 0};
    int64_t ntet = 0;
    long long *bin_start = nullptr;    // kBins^3 + 1
    int *bin_list = nullptr;
    // preprocessed query points
    bool pre = false;
    int64_t pre_n = 0;
    int *pre_idx = nullptr;            // nearest: n ids; linear: 4 n ids
    float *pre_w = nullptr;            // linear: 4 n weights
};

namespace ffdas {
namespace {

constexpr int kBins = 128;
constexpr int64_t kBins3 = (int64_t)kBins * kBins * kBins;

// the six tetrahedra of a cell as corner codes (bit 0: +x, bit 1: +y, bit 2: +z), in the
// reference's order (src/interpolation/spatial_hash_kernels.cuh:183-190): the split fixes
// which vertices a point is interpolated from
__constant__ unsigned char kTet[6][4] = {{0, 1, 2, 4}, {1, 3, 2, 7}, {1, 2, 4, 7},
                                         {2, 4, 6, 7}, {1, 4, 5, 7}, {4, 6, 5, 7}};

__device__ __forceinline__ int4 tet_vertices(long long tid, int nx, int ny) {
    const long long cell = tid / 6;
    const int t = (int)(tid - cell * 6);
    const long long cx = cell % (nx - 1), cy = (cell / (nx - 1)) % (ny - 1),
                    cz = cell / ((long long)(nx - 1) * (ny - 1));
    const long long base = cx + cy * nx + cz * (long long)ny * nx;
    int v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int code = kTet[t][c];
        v[c] = (int)(base + (code & 1) + ((code >> 1) & 1) * nx + ((code >> 2) & 1) * (long long)ny * nx);
    }
    return int4{v[0], v[1], v[2], v[3]};
}

// A tetrahedron in the bin lists: key = base vertex * 8 + t, base = cx + cy nx + cz ny nx of
// its cell.  Keys order exactly as the ids cell * 6 + t do (cells in (cz, cy, cx) order, then
// t), so the lowest passing id is the lowest passing key, and the vertices follow from the
// key without divisions.
__device__ __forceinline__ int tet_key(long long tid, int nx, int ny) {
    const long long cell = tid / 6;
    const int t = (int)(tid - cell * 6);
    const long long cx = cell % (nx - 1), cy = (cell / (nx - 1)) % (ny - 1),
                    cz = cell / ((long long)(nx - 1) * (ny - 1));
    return (int)((cx + cy * nx + cz * (long long)ny * nx) * 8 + t);
}

__device__ __forceinline__ int4 key_vertices(int key, int nx, int ny) {
    const int base = key >> 3, t = key & 7;
    int v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int code = kTet[t][c];
        v[c] = base + (code & 1) + ((code >> 1) & 1) * nx + ((code >> 2) & 1) * ny * nx;
    }
    return int4{v[0], v[1], v[2], v[3]};
}

struct Box {
    float mn[3], sc[3];
};

// round-to-nearest bin of (p - min) * scale with negatives saturated to 0 (NaN counts as
// outside); returns false past the last bin
__device__ __forceinline__ bool quantize(const Box &b, float x, float y, float z, int &ix, int &iy,
                                         int &iz) {
    const float q[3] = {(x - b.mn[0]) * b.sc[0], (y - b.mn[1]) * b.sc[1], (z - b.mn[2]) * b.sc[2]};
    int r[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (!(q[a] <= (float)(kBins - 1) + 0.5f)) return false;   // beyond the range, or NaN
        const float v = rintf(q[a] > 0.0f ? q[a] : 0.0f);
        if (v > (float)(kBins - 1)) return false;
        r[a] = (int)v;
    }
    ix = r[0];
    iy = r[1];
    iz = r[2];
    return true;
}

__global__ __launch_bounds__(256) void interp_bbox_kernel(const float *__restrict__ v, int64_t n,
                                                          float *__restrict__ part) {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float x = v[3 * i + a];
            mn[a] = fminf(mn[a], x);
            mx[a] = fmaxf(mx[a], x);
        }
    }
    __shared__ float red[6][4];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        for (int o = 32; o > 0; o >>= 1) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], o));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o));
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            red[a][w] = mn[a];
            red[3 + a][w] = mx[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        float r = red[a][0];
        for (int k = 1; k < 4; k++) r = a < 3 ? fminf(r, red[a][k]) : fmaxf(r, red[a][k]);
        part[blockIdx.x * 6 + a] = r;
    }
}

__global__ void interp_zero_kernel(long long *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// quantised bounding box of tetrahedron tid, clamped to the bins
__device__ __forceinline__ void tet_bins(const Box &b, const float *__restrict__ v, long long tid,
                                         int nx, int ny, int lo[3], int hi[3]) {
    const int4 id = tet_vertices(tid, nx, ny);
    const int ids[4] = {id.x, id.y, id.z, id.w};
    lo[0] = lo[1] = lo[2] = kBins - 1;
    hi[0] = hi[1] = hi[2] = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float *p = v + 3 * (int64_t)ids[c];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            // vertices lie inside the box: (p - min) * scale is in [0, 127] up to rounding
            const float q = rintf(fmaxf((p[a] - b.mn[a]) * b.sc[a], 0.0f));
            const int k = (int)fminf(q, (float)(kBins - 1));
            lo[a] = min(lo[a], k);
            hi[a] = max(hi[a], k);
        }
    }
}

__global__ __launch_bounds__(256) void interp_count_kernel(Box b, const float *__restrict__ v,
                                                           int64_t ntet, int nx, int ny,
                                                           long long *__restrict__ cnt) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= ntet) return;
    int lo[3], hi[3];
    tet_bins(b, v, t, nx, ny, lo, hi);
    for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int x = lo[0]; x <= hi[0]; x++)
                atomicAdd((unsigned long long *)cnt + (x + kBins * (y + kBins * z)), 1ull);
}

__global__ __launch_bounds__(256) void interp_fill_kernel(Box b, const float *__restrict__ v,
                                                          int64_t ntet, int nx, int ny,
                                                          const long long *__restrict__ start,
                                                          long long *__restrict__ cursor,
                                                          int *__restrict__ list) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= ntet) return;
    int lo[3], hi[3];
    tet_bins(b, v, t, nx, ny, lo, hi);
    for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int x = lo[0]; x <= hi[0]; x++) {
                const int64_t bin = x + kBins * (y + kBins * z);
                const long long k = (long long)atomicAdd((unsigned long long *)cursor + bin, 1ull);
                list[start[bin] + k] = tet_key(t, nx, ny);
            }
}

// Workgroups are dealt to the 8 XCDs round-robin by linear id.  This bijection hands each
// XCD groups of kXcdGroup consecutive (query block, chunk) tiles instead of single ones, so
// neighbouring query blocks, which share vertices, are served by one L2, while the XCDs
// still take turns over the whole grid (the located points are unevenly spread: whole
// contiguous eighths would leave some XCDs with most of the work).  Tiles past the last
// full round of 8 groups keep their own id.  Returns the tile as (x, y).
constexpr int kXcdGroup = 8;

__device__ __forceinline__ void xcd_tile(int &bx, int &by) {
    const int total = gridDim.x * gridDim.y;
    const int lin = blockIdx.y * gridDim.x + blockIdx.x;
    constexpr int R = 8 * kXcdGroup;
    int L = lin;
    if (lin < total / R * R) {
        const int base = lin / R * R, o = lin - base;   // o = j * 8 + xcd within the round
        L = base + (o % 8) * kXcdGroup + o / 8;
    }
    bx = L % gridDim.x;
    by = L / gridDim.x;
}

// FP32 barycentric coordinates of p in (v0, v1, v2, v3) and whether all are >= 0, in the
// reference's operation order (spatial_hash_kernels.cuh:62-90), with no FMA contraction
__device__ bool barycentric(const float *p, const float *v0, const float *v1, const float *v2,
                            const float *v3, float w[4]) {
#pragma clang fp contract(off)
    const float a0 = v1[0] - v0[0], a1 = v1[1] - v0[1], a2 = v1[2] - v0[2];
    const float b0 = v2[0] - v0[0], b1 = v2[1] - v0[1], b2 = v2[2] - v0[2];
    const float c0 = v3[0] - v0[0], c1 = v3[1] - v0[1], c2 = v3[2] - v0[2];
    const float q0 = p[0] - v0[0], q1 = p[1] - v0[1], q2 = p[2] - v0[2];
    // cross(b, c), dot(a, .)
    const float bc0 = b1 * c2 - b2 * c1, bc1 = b2 * c0 - b0 * c2, bc2 = b0 * c1 - b1 * c0;
    const float denom = a0 * bc0 + a1 * bc1 + a2 * bc2;
    if (fabsf(denom) < 1e-12f) return false;
    const float qc0 = q1 * c2 - q2 * c1, qc1 = q2 * c0 - q0 * c2, qc2 = q0 * c1 - q1 * c0;
    const float bq0 = b1 * q2 - b2 * q1, bq1 = b2 * q0 - b0 * q2, bq2 = b0 * q1 - b1 * q0;
    const float n1 = q0 * bc0 + q1 * bc1 + q2 * bc2;
    const float n2 = a0 * qc0 + a1 * qc1 + a2 * qc2;
    const float n3 = a0 * bq0 + a1 * bq1 + a2 * bq2;
    // most candidates fail on a sign: a numerator of the other sign than denom with
    // |n| >= 1e-30 |denom| gives a quotient below -1e-30 (never a rounded -0), which the
    // test below rejects, so the divisions are skipped for it
    const float lim = 1e-30f * fabsf(denom);
    const bool neg = denom < 0.0f;
    if (((n1 < 0.0f) != neg && n1 != 0.0f && fabsf(n1) >= lim) ||
        ((n2 < 0.0f) != neg && n2 != 0.0f && fabsf(n2) >= lim) ||
        ((n3 < 0.0f) != neg && n3 != 0.0f && fabsf(n3) >= lim))
        return false;
    w[1] = n1 / denom;
    w[2] = n2 / denom;
    w[3] = n3 / denom;
    w[0] = 1.0f - w[1] - w[2] - w[3];
    return w[0] >= 0.0f && w[1] >= 0.0f && w[2] >= 0.0f && w[3] >= 0.0f;
}

// mode 0: idx[i] = vertex of largest weight (the reference's tie order); mode 1: idx[4 i..]
// and w[4 i..]; -1 where no tetrahedron holds the point
__global__ __launch_bounds__(256) void interp_locate_kernel(
    Box b, const float *__restrict__ v, int nx, int ny, int64_t ntet,
    const long long *__restrict__ start, const int *__restrict__ list,
    const float *__restrict__ qp, int64_t n, int mode, int *__restrict__ idx,
    float *__restrict__ wout) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float p[3] = {qp[3 * i], qp[3 * i + 1], qp[3 * i + 2]};
    int ix, iy, iz;
    int best = INT_MAX;
    int4 bid{-1, -1, -1, -1};
    float bw[4] = {0, 0, 0, 0};
    if (ntet > 0 && quantize(b, p[0], p[1], p[2], ix, iy, iz)) {
        const int64_t bin = ix + kBins * (iy + kBins * (int64_t)iz);
        for (long long e = start[bin]; e < start[bin + 1]; e++) {
            const int t = list[e];
            if (t >= best) continue;
            const int4 id = key_vertices(t, nx, ny);
            float w[4];
            if (barycentric(p, v + 3 * (int64_t)id.x, v + 3 * (int64_t)id.y, v + 3 * (int64_t)id.z,
                            v + 3 * (int64_t)id.w, w)) {
                best = t;
                bid = id;
#pragma unroll
                for (int c = 0; c < 4; c++) bw[c] = w[c];
            }
        }
    }
    if (mode == FFDAS_INTERP_NEAREST) {
        int r = -1;
        if (best != INT_MAX) {   // float4_max_by_key (spatial_hash_kernels.cuh:51-59)
            const int ixy = bw[0] > bw[1] ? bid.x : bid.y;
            const float mxy = ixy == bid.x ? bw[0] : bw[1];
            const int izw = bw[2] > bw[3] ? bid.z : bid.w;
            const float mzw = izw == bid.z ? bw[2] : bw[3];
            r = mxy > mzw ? ixy : izw;
        }
        idx[i] = r;
    } else {
        reinterpret_cast<int4 *>(idx)[i] = bid;
        reinterpret_cast<float4 *>(wout)[i] = float4{bw[0], bw[1], bw[2], bw[3]};
    }
}

// ---------------------------------------------------------------------------------------
// Location by bins: the query points are counted and ordered by bin, and one wave per bin
// loads the bin's points (one per lane) and candidate tetrahedra (key, vertices and bounding
// box, one per lane and chunk) once and tests every point of the bin against them, so each
// candidate's vertices are fetched once per bin rather than once per point, and no load waits
// inside the per-point loop (the per-point scan re-read about 3.2 KB per point on a 512 x 96 x 96
// sector with 256^3 query points).  The result is the same: per point, the lowest passing key
// of its bin.
// ---------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void interp_pad_kernel(const float *__restrict__ v, int64_t n,
                                                         float4 *__restrict__ v4) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v4[i] = float4{v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.f};
}

__global__ void interp_zero_i32_kernel(int *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// "no tetrahedron holds the point" in the lookup of point i
__device__ __forceinline__ void write_missing(int mode, int64_t i, int *idx, float *wout) {
    if (mode == FFDAS_INTERP_NEAREST) {
        idx[i] = -1;
    } else {
        reinterpret_cast<int4 *>(idx)[i] = int4{-1, -1, -1, -1};
        reinterpret_cast<float4 *>(wout)[i] = float4{0.f, 0.f, 0.f, 0.f};
    }
}

// pbin[i] = the point's bin, or -1 (outside the bins, or no tetrahedra: its lookup is written
// here); cnt[bin] counts the points of each bin
__global__ __launch_bounds__(256) void interp_point_bin_kernel(
    Box b, const float *__restrict__ qp, int64_t n, int64_t ntet, int mode,
    int *__restrict__ pbin, int *__restrict__ cnt, int *__restrict__ idx,
    float *__restrict__ wout) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int ix, iy, iz, bin = -1;
    if (ntet > 0 && quantize(b, qp[3 * i], qp[3 * i + 1], qp[3 * i + 2], ix, iy, iz))
        bin = ix + kBins * (iy + kBins * iz);
    pbin[i] = bin;
    if (bin >= 0)
        atomicAdd(cnt + bin, 1);
    else
        write_missing(mode, i, idx, wout);
}

// order[start[bin] ..) = the points of each bin (counting cnt back down to 0)
__global__ __launch_bounds__(256) void interp_point_order_kernel(const int *__restrict__ pbin,
                                                                 int64_t n,
                                                                 const int *__restrict__ start,
                                                                 int *__restrict__ cnt,
                                                                 int *__restrict__ order) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int bin = pbin[i];
    if (bin < 0) return;
    order[start[bin] + atomicSub(cnt + bin, 1) - 1] = (int)i;
}

// One wave per bin.  The bin's points come in batches of 64, one per lane (its id and
// position, loaded once); the bin's candidates in groups of up to two chunks of 64, lane c
// holding candidate c of each chunk in registers (its key, vertices and bounding box), each
// group loaded once per batch.  For each point t of the batch (read from lane t), a chunk is tested
// only if the point lies in the box of one of its candidates, widened by 5 % of its extent plus
// 1e-5 of the coordinates' magnitude per axis (the FP32 barycentric weights of a tetrahedron
// that the 1e-12 volume test admits err far less than that, so no candidate the test would
// pass is skipped); then every lane runs the reference's FP32 test on its candidate.  A chunk's passing lanes come from one ballot, usually none or one, whose keys a scalar
// loop compares with point t's best so far; point t's best key and the winner's weights are
// kept in lane t.  Lane t then writes the lookup of point t (mode 0: the vertex of largest
// weight; mode 1: ids and weights), or "none".
constexpr int kLocateChunks = 2;

__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}


__global__ __launch_bounds__(256) void interp_locate_bin_kernel(
    const float4 *__restrict__ v4, int nx, int ny, const long long *__restrict__ start,
    const int *__restrict__ list, const float *__restrict__ qp, const int *__restrict__ pstart,
    const int *__restrict__ order, int mode, int *__restrict__ idx, float *__restrict__ wout) {
    const int lane = threadIdx.x & 63;
    const int64_t bin = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (bin >= kBins3) return;
    const int ps = pstart[bin], pe = pstart[bin + 1];
    if (ps == pe) return;
    const long long cs = start[bin], ce = start[bin + 1];
    constexpr int G = 64 * kLocateChunks;
    int key[kLocateChunks];
    float lo[kLocateChunks][3], hi[kLocateChunks][3], vp[kLocateChunks][4][3];
    // the group of chunks from g0: key INT_MAX (and an empty box) past the list
    const auto load = [&](long long g0) {
#pragma unroll
        for (int c = 0; c < kLocateChunks; ++c) {
            const long long e = g0 + 64 * c + lane;
            key[c] = INT_MAX;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[c][a] = 3.4e38f;
                hi[c][a] = -3.4e38f;
            }
            if (e < ce) {
                key[c] = list[e];
                const int4 id = key_vertices(key[c], nx, ny);
                const int ids[4] = {id.x, id.y, id.z, id.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 x = v4[ids[q]];
                    vp[c][q][0] = x.x;
                    vp[c][q][1] = x.y;
                    vp[c][q][2] = x.z;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        lo[c][a] = fminf(lo[c][a], vp[c][q][a]);
                        hi[c][a] = fmaxf(hi[c][a], vp[c][q][a]);
                    }
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float m = 0.05f * (hi[c][a] - lo[c][a]) +
                                    1e-5f * (fabsf(lo[c][a]) + fabsf(hi[c][a])) + 1e-30f;
                    lo[c][a] -= m;
                    hi[c][a] += m;
                }
            }
        }
    };
    const bool one_group = ce - cs <= G;
    if (one_group) load(cs);
    for (int q0 = ps; q0 < pe; q0 += 64) {
        const int np = min(64, pe - q0);
        // point q0 + lane: its id, position, and best key and weights so far
        int mi = -1, mbest = INT_MAX;
        float mp[3] = {0.f, 0.f, 0.f}, mw[4] = {0.f, 0.f, 0.f, 0.f};
        if (lane < np) {
            mi = order[q0 + lane];
#pragma unroll
            for (int a = 0; a < 3; ++a) mp[a] = qp[3 * (int64_t)mi + a];
        }
        for (long long g0 = cs; g0 < ce; g0 += G) {
            if (!one_group) load(g0);
            for (int t = 0; t < np; ++t) {
                const float p[3] = {readlane_f(mp[0], t), readlane_f(mp[1], t),
                                    readlane_f(mp[2], t)};
                int best = __builtin_amdgcn_readlane(mbest, t);
                int won = -1;   // uniform
#pragma unroll
                for (int c = 0; c < kLocateChunks; ++c) {
                    if (g0 + 64 * c >= ce) break;   // uniform
                    const bool near = key[c] < best && p[0] >= lo[c][0] && p[0] <= hi[c][0] &&
                                      p[1] >= lo[c][1] && p[1] <= hi[c][1] &&
                                      p[2] >= lo[c][2] && p[2] <= hi[c][2];
                    if (__ballot(near) == 0) continue;   // no candidate of the chunk can pass
                    float w[4] = {0.f, 0.f, 0.f, 0.f};
                    bool pass = false;
                    if (key[c] < best)
                        pass = barycentric(p, vp[c][0], vp[c][1], vp[c][2], vp[c][3], w);
                    uint64_t m = __ballot(pass);
                    int cw = -1;
                    while (m) {
                        const int l = __builtin_ctzll(m);
                        m &= m - 1;
                        const int k = __builtin_amdgcn_readlane(key[c], l);
                        if (k < best) {
                            best = k;
                            cw = l;
                        }
                    }
                    if (cw >= 0) {
                        won = cw;
                        // the winner's weights into lane t
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const float wu = readlane_f(w[u], cw);
                            mw[u] = lane == t ? wu : mw[u];
                        }
                    }
                }
                if (won >= 0 && lane == t) mbest = best;
            }
        }
        if (lane < np) {
            if (mbest == INT_MAX) {
                write_missing(mode, mi, idx, wout);
            } else {
                const int4 bid = key_vertices(mbest, nx, ny);
                if (mode == FFDAS_INTERP_NEAREST) {
                    // float4_max_by_key (spatial_hash_kernels.cuh:51-59)
                    const int ixy = mw[0] > mw[1] ? bid.x : bid.y;
                    const float mxy = ixy == bid.x ? mw[0] : mw[1];
                    const int izw = mw[2] > mw[3] ? bid.z : bid.w;
                    const float mzw = izw == bid.z ? mw[2] : mw[3];
                    idx[mi] = mxy > mzw ? ixy : izw;
                } else {
                    reinterpret_cast<int4 *>(idx)[mi] = bid;
                    reinterpret_cast<float4 *>(wout)[mi] = float4{mw[0], mw[1], mw[2], mw[3]};
                }
            }
        }
    }
}

// element types: real or complex (two components), FP32 or FP64
template <typename R, int CW>
struct El {
    R c[CW];
};

// Outputs are written once and never read back by the kernel: streamed past the caches so
// they do not evict the vertex blocks that neighbouring query points gather.
template <typename R, int CW>
__device__ __forceinline__ void put_out(El<R, CW> *p, const El<R, CW> &v) {
#pragma unroll
    for (int c = 0; c < CW; c++) __builtin_nontemporal_store(v.c[c], &p->c[c]);
}

constexpr int kFramesPerThread = 8;   // frames per chunk: the smallest, for the extent check

// frames per chunk: 16 for 4-byte elements (a 64-byte vector load per vertex), else 8
template <typename R, int CW>
constexpr int frames_per_chunk() {
    return sizeof(R) * CW == 4 ? 16 : 8;
}

// the frames of one vertex in one chunk, contiguous: one vector load per vertex
template <typename R, int CW>
struct alignas(16) Frames {
    El<R, CW> f[frames_per_chunk<R, CW>()];
};

// x [frame][vertex] -> xfi [chunk][vertex][frame in chunk], zero past the last frame; each
// thread reads its vertex from the chunk's frames (coalesced per frame) and writes the
// vertex's frames as one contiguous block (coalesced across threads)
template <typename R, int CW>
__global__ __launch_bounds__(256) void interp_frames_kernel(const El<R, CW> *__restrict__ x,
                                                            Frames<R, CW> *__restrict__ xfi,
                                                            int64_t npts, int64_t B) {
    constexpr int FPC = frames_per_chunk<R, CW>();
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= npts) return;
    const int64_t b0 = (int64_t)blockIdx.y * FPC;
    Frames<R, CW> o;
#pragma unroll
    for (int f = 0; f < FPC; f++) {
        if (b0 + f < B) {
            o.f[f] = x[(b0 + f) * npts + v];
        } else {
#pragma unroll
            for (int c = 0; c < CW; c++) o.f[f].c[c] = R(0);
        }
    }
    xfi[blockIdx.y * npts + v] = o;
}

template <typename R, int CW>
__global__ __launch_bounds__(256) void interp_nearest_kernel(const Frames<R, CW> *__restrict__ xfi,
                                                             El<R, CW> *__restrict__ out,
                                                             const int *__restrict__ idx, int64_t n,
                                                             int64_t npts, int64_t B,
                                                             El<R, CW> fill) {
    constexpr int FPC = frames_per_chunk<R, CW>();
    int bx, by;
    xcd_tile(bx, by);
    const int64_t i = (int64_t)bx * 256 + threadIdx.x;
    if (i >= n) return;
    const int k = idx[i];
    const int64_t b0 = (int64_t)by * FPC;
    const int nf = (int)(B - b0 < FPC ? B - b0 : FPC);
    if (k >= 0) {
        const Frames<R, CW> v = xfi[by * npts + k];
#pragma unroll
        for (int f = 0; f < FPC; f++)
            if (f < nf) put_out(out + (b0 + f) * n + i, v.f[f]);
    } else {
#pragma unroll
        for (int f = 0; f < FPC; f++)
            if (f < nf) put_out(out + (b0 + f) * n + i, fill);
    }
}

template <typename R, int CW>
__global__ __launch_bounds__(256) void interp_linear_kernel(const Frames<R, CW> *__restrict__ xfi,
                                                            El<R, CW> *__restrict__ out,
                                                            const int4 *__restrict__ idx,
                                                            const float4 *__restrict__ wt, int64_t n,
                                                            int64_t npts, int64_t B,
                                                            El<R, CW> fill) {
#pragma clang fp contract(off)
    constexpr int FPC = frames_per_chunk<R, CW>();
    int bx, by;
    xcd_tile(bx, by);
    const int64_t i = (int64_t)bx * 256 + threadIdx.x;
    if (i >= n) return;
    // the lookup is read once per frame chunk: streamed, like the outputs
    typedef int I4 __attribute__((ext_vector_type(4)));
    typedef float F4 __attribute__((ext_vector_type(4)));
    const I4 kv = __builtin_nontemporal_load(reinterpret_cast<const I4 *>(idx) + i);
    const F4 wv = __builtin_nontemporal_load(reinterpret_cast<const F4 *>(wt) + i);
    const int4 k = make_int4(kv.x, kv.y, kv.z, kv.w);
    const float4 w = make_float4(wv.x, wv.y, wv.z, wv.w);
    const int64_t b0 = (int64_t)by * FPC;
    const int nf = (int)(B - b0 < FPC ? B - b0 : FPC);
    const Frames<R, CW> *xc = xfi + by * npts;
    // unlocated points read vertex 0 (always present) and write the fill value: the loads
    // stay unconditional, so the four blocks live in registers
    const bool in = k.x >= 0;
    const Frames<R, CW> v0 = xc[in ? k.x : 0], v1 = xc[in ? k.y : 0], v2 = xc[in ? k.z : 0],
                        v3 = xc[in ? k.w : 0];
#pragma unroll
    for (int f = 0; f < FPC; f++) {
        El<R, CW> r;
        // w0 v0 + w1 v1 + w2 v2 + w3 v3, left to right (interpolation_kernels.cuh:110-115)
#pragma unroll
        for (int c = 0; c < CW; c++)
            r.c[c] = in ? (R)w.x * v0.f[f].c[c] + (R)w.y * v1.f[f].c[c] + (R)w.z * v2.f[f].c[c] +
                              (R)w.w * v3.f[f].c[c]
                        : fill.c[c];
        if (f < nf) put_out(out + (b0 + f) * n + i, r);
    }
}

void free_plan_device(ffdas_interpolation_plan *p) {
    for (void *q : {(void *)p->verts, (void *)p->verts4, (void *)p->bin_start, (void *)p->bin_list,
                    (void *)p->pre_idx, (void *)p->pre_w})
        if (q) (void)hipFree(q);
    p->verts = nullptr;
    p->verts4 = nullptr;
    p->bin_start = nullptr;
    p->bin_list = nullptr;
    p->pre_idx = nullptr;
    p->pre_w = nullptr;
}

Box box_of(const ffdas_interpolation_plan &p) {
    Box b;
    for (int a = 0; a < 3; a++) {
        b.mn[a] = p.bmin[a];
        b.sc[a] = p.scale[a];
    }
    return b;
}

int grid1(int64_t n) { return (int)std::max<int64_t>(1, (n + 255) / 256); }

// locate n points into idx (and w): device buffers sized by the caller.  By bins (the points
// counted, scanned and ordered by bin in the handle's workspace, then one wave per bin) when
// the point ids fit an int and FFDAS_INTERP_LOCATE is not "point"; otherwise one thread per
// point scanning its bin.
ffdas_error_t locate(ffdas_context &h, const ffdas_interpolation_plan &p, const float *qp, int64_t n,
                     int *idx, float *w) {
    const char *how = std::getenv("FFDAS_INTERP_LOCATE");
    const bool by_bins = n < INT_MAX && !(how && std::string(how) == "point");
    if (!by_bins) {
        ProfScope ps(h.stream, "interp_locate");
        hipLaunchKernelGGL(interp_locate_kernel, dim3(grid1(n)), dim3(256), 0, h.stream,
                           box_of(p), p.verts, (int)p.nx, (int)p.ny, p.ntet, p.bin_start,
                           p.bin_list, qp, n, p.mode, idx, w);
        return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
    }
    hipStream_t st = h.stream;
    const size_t nb = (size_t)kBins3 + 1;
    size_t scan_bytes = 0;
    int *no_int = nullptr;
    if (rocprim::exclusive_scan(nullptr, scan_bytes, no_int, no_int, 0, nb, rocprim::plus<int>(),
                                st) != hipSuccess)
        return FFDAS_ERROR_CUDA_LAUNCH_FAILED;
    const auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_pbin = up((size_t)n * 4), b_cnt = up(nb * 4), b_scan = up(scan_bytes);
    void *ws = nullptr;
    FFDAS_TRY(h.workspace(2 * b_pbin + 2 * b_cnt + b_scan, &ws));
    h.xcopy.valid = false;   // the converted-x record of the delay-and-sum calls
    char *c = (char *)ws;
    int *pbin = (int *)c, *order = (int *)(c + b_pbin);
    int *cnt = (int *)(c + 2 * b_pbin), *pstart = (int *)(c + 2 * b_pbin + b_cnt);
    void *scan_tmp = c + 2 * b_pbin + 2 * b_cnt;
    ProfScope ps(st, "interp_locate");
    hipLaunchKernelGGL(interp_zero_i32_kernel, dim3(1024), dim3(256), 0, st, cnt, (int64_t)nb);
    hipLaunchKernelGGL(interp_point_bin_kernel, dim3(grid1(n)), dim3(256), 0, st, box_of(p), qp, n,
                       p.ntet, p.mode, pbin, cnt, idx, w);
    if (rocprim::exclusive_scan(scan_tmp, scan_bytes, cnt, pstart, 0, nb, rocprim::plus<int>(),
                                st) != hipSuccess)
        return FFDAS_ERROR_CUDA_LAUNCH_FAILED;
    hipLaunchKernelGGL(interp_point_order_kernel, dim3(grid1(n)), dim3(256), 0, st, pbin, n,
                       pstart, cnt, order);
    hipLaunchKernelGGL(interp_locate_bin_kernel, dim3((unsigned)((kBins3 + 3) / 4)), dim3(256), 0,
                       st, p.verts4, (int)p.nx, (int)p.ny, p.bin_start, p.bin_list, qp, pstart,
                       order, p.mode, idx, w);
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

ffdas_error_t alloc_lookup(int mode, int64_t n, int **idx, float **w) {
    const size_t ni = (size_t)n * (mode == FFDAS_INTERP_NEAREST ? 1 : 4);
    if (hipMalloc((void **)idx, ni * sizeof(int)) != hipSuccess) return FFDAS_ERROR_ALLOCATION_FAILED;
    *w = nullptr;
    if (mode != FFDAS_INTERP_NEAREST &&
        hipMalloc((void **)w, (size_t)n * 4 * sizeof(float)) != hipSuccess) {
        (void)hipFree(*idx);
        *idx = nullptr;
        return FFDAS_ERROR_ALLOCATION_FAILED;
    }
    return FFDAS_SUCCESS;
}

template <typename R, int CW>
ffdas_error_t evaluate(ffdas_context &h, int mode, const void *x, void *out, const int *idx,
                       const float *w, int64_t n, int64_t B, int64_t npts, const void *fill) {
    using E = El<R, CW>;
    using F = Frames<R, CW>;
    E f;
    for (int c = 0; c < CW; c++) f.c[c] = static_cast<const R *>(fill)[c];
    constexpr int FPC = frames_per_chunk<R, CW>();
    const int64_t chunks = (B + FPC - 1) / FPC;
    // the frame-inner copy of x lives in the handle's workspace (which then no longer holds
    // the converted x of an earlier delay-and-sum call)
    void *ws = nullptr;
    h.xcopy.valid = false;
    FFDAS_TRY(h.workspace((size_t)chunks * npts * sizeof(F), &ws));
    F *xfi = static_cast<F *>(ws);
    {
        ProfScope ps(h.stream, "interp_frames");
        hipLaunchKernelGGL((interp_frames_kernel<R, CW>), dim3(grid1(npts), (unsigned)chunks),
                           dim3(256), 0, h.stream, static_cast<const E *>(x), xfi, npts, B);
    }
    const dim3 grid(grid1(n), (unsigned)chunks);
    ProfScope ps(h.stream, "interp_eval");
    if (mode == FFDAS_INTERP_NEAREST)
        hipLaunchKernelGGL((interp_nearest_kernel<R, CW>), grid, dim3(256), 0, h.stream, xfi,
                           static_cast<E *>(out), idx, n, npts, B, f);
    else
        hipLaunchKernelGGL((interp_linear_kernel<R, CW>), grid, dim3(256), 0, h.stream, xfi,
                           static_cast<E *>(out), reinterpret_cast<const int4 *>(idx),
                           reinterpret_cast<const float4 *>(w), n, npts, B, f);
    return hip_status(hipGetLastError(), FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

ffdas_error_t build_plan(ffdas_context &h, ffdas_interpolation_plan &p, const float *gridpos) {
    const int64_t npts = p.nx * p.ny * p.nz;
    hipStream_t st = h.stream;
    if (hipMalloc((void **)&p.verts, (size_t)npts * 3 * sizeof(float)) != hipSuccess)
        return FFDAS_ERROR_ALLOCATION_FAILED;
    FFDAS_TRY(hip_status(hipMemcpyAsync(p.verts, gridpos, (size_t)npts * 3 * sizeof(float),
                                        hipMemcpyDeviceToDevice, st),
                         FFDAS_ERROR_CUDA_MEMCPY_FAILED));
    if (hipMalloc((void **)&p.verts4, (size_t)npts * sizeof(float4)) != hipSuccess)
        return FFDAS_ERROR_ALLOCATION_FAILED;
    hipLaunchKernelGGL(interp_pad_kernel, dim3(grid1(npts)), dim3(256), 0, st, p.verts, npts,
                       p.verts4);
    // bounding box: per-block partials, finished on the host
    const int nblk = (int)std::min<int64_t>(256, std::max<int64_t>(1, (npts + 255) / 256));
    float *part = nullptr;
    if (hipMalloc((void **)&part, (size_t)nblk * 6 * sizeof(float)) != hipSuccess)
        return FFDAS_ERROR_ALLOCATION_FAILED;
    hipLaunchKernelGGL(interp_bbox_kernel, dim3(nblk), dim3(256), 0, st, p.verts, npts, part);
    std::vector<float> hp((size_t)nblk * 6);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(part);
    FFDAS_TRY(hip_status(e));
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int k = 0; k < nblk; k++)
        for (int a = 0; a < 3; a++) {
            mn[a] = std::min(mn[a], hp[k * 6 + a]);
            mx[a] = std::max(mx[a], hp[k * 6 + 3 + a]);
        }
    for (int a = 0; a < 3; a++) {
        p.bmin[a] = mn[a];
        // the reference's scale, in FP32 (spatial_hash.cu:51-55)
        volatile float ext = mx[a] - mn[a];
        p.scale[a] = (float)(kBins - 1) / ext;
    }
    p.ntet = 6 * (p.nx - 1) * (p.ny - 1) * (p.nz - 1);
    if (hipMalloc((void **)&p.bin_start, (size_t)(kBins3 + 1) * sizeof(long long)) != hipSuccess)
        return FFDAS_ERROR_ALLOCATION_FAILED;
    long long *cnt = nullptr;
    if (hipMalloc((void **)&cnt, (size_t)(kBins3 + 1) * sizeof(long long)) != hipSuccess)
        return FFDAS_ERROR_ALLOCATION_FAILED;
    const Box b = box_of(p);
    const int zg = (int)std::min<int64_t>(4096, (kBins3 + 1 + 255) / 256);
    hipLaunchKernelGGL(interp_zero_kernel, dim3(zg), dim3(256), 0, st, cnt, kBins3 + 1);
    if (p.ntet > 0)
        hipLaunchKernelGGL(interp_count_kernel, dim3(grid1(p.ntet)), dim3(256), 0, st, b, p.verts,
                           p.ntet, (int)p.nx, (int)p.ny, cnt);
    // exclusive scan of the counts (the sentinel count is 0, so start[kBins3] = total)
    size_t tmp_bytes = 0;
    e = rocprim::exclusive_scan(nullptr, tmp_bytes, cnt, p.bin_start, 0ll, (size_t)(kBins3 + 1),
                                rocprim::plus<long long>(), st);
    void *tmp = nullptr;
    if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes);
    if (e == hipSuccess)
        e = rocprim::exclusive_scan(tmp, tmp_bytes, cnt, p.bin_start, 0ll, (size_t)(kBins3 + 1),
                                    rocprim::plus<long long>(), st);
    long long total = 0;
    if (e == hipSuccess)
        e = hipMemcpyAsync(&total, p.bin_start + kBins3, sizeof(long long), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess) {
        (void)hipFree(cnt);
        return hip_status(e);
    }
    if (total > INT_MAX || p.ntet > INT_MAX) {
        (void)hipFree(cnt);
        return FFDAS_ERROR_DIMS_TOO_LARGE;
    }
    if (hipMalloc((void **)&p.bin_list, (size_t)std::max<long long>(total, 1) * sizeof(int)) != hipSuccess) {
        (void)hipFree(cnt);
        return FFDAS_ERROR_ALLOCATION_FAILED;
    }
    hipLaunchKernelGGL(interp_zero_kernel, dim3(zg), dim3(256), 0, st, cnt, kBins3 + 1);
    if (p.ntet > 0)
        hipLaunchKernelGGL(interp_fill_kernel, dim3(grid1(p.ntet)), dim3(256), 0, st, b, p.verts,
                           p.ntet, (int)p.nx, (int)p.ny, p.bin_start, cnt, p.bin_list);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);   // cnt is freed next
    (void)hipFree(cnt);
    return hip_status(e, FFDAS_ERROR_CUDA_LAUNCH_FAILED);
}

}  // namespace
}  // namespace ffdas

using namespace ffdas;

static ffdas_error_t create_interpolation_plan_call(ffdas_handle_t handle,
                                                         ffdas_interpolation_plan_t *plan,
                                                         int64_t nx, int64_t ny, int64_t nz,
                                                         const float *gridpos,
                                                         ffdas_interp_mode_t mode) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    if (!plan || !gridpos) return FFDAS_ERROR_NULL_POINTER;
    *plan = nullptr;
    if (nx <= 0 || ny <= 0 || nz <= 0) return FFDAS_ERROR_INVALID_DIMS;
    if (mode != FFDAS_INTERP_NEAREST && mode != FFDAS_INTERP_LINEAR)
        return FFDAS_ERROR_INVALID_ARGUMENT;
    // vertex ids are 32-bit in the lookup arrays; the bin lists hold vertex * 8 + t
    if (nx > INT_MAX / 8 / ny / nz || 6 * (nx - 1) * (ny - 1) * (nz - 1) > INT_MAX)
        return FFDAS_ERROR_DIMS_TOO_LARGE;
    FFDAS_TRY(handle->check_device());
    auto *p = new (std::nothrow) ffdas_interpolation_plan;
    if (!p) return FFDAS_ERROR_HOST_MEMORY_INSUFFICIENT;
    p->device = handle->device_id;
    p->nx = nx;
    p->ny = ny;
    p->nz = nz;
    p->mode = mode;
    const ffdas_error_t err = build_plan(*handle, *p, gridpos);
    if (err != FFDAS_SUCCESS) {
        (void)hipStreamSynchronize(handle->stream);
        free_plan_device(p);
        delete p;
        return err;
    }
    *plan = p;
    return FFDAS_SUCCESS;
}

extern "C" ffdas_error_t ffdas_create_interpolation_plan(ffdas_handle_t handle,
                                                         ffdas_interpolation_plan_t *plan,
                                                         int64_t nx, int64_t ny, int64_t nz,
                                                         const float *gridpos,
                                                         ffdas_interp_mode_t mode) {
    ffdas::CallScope cs(handle, "ffdas_create_interpolation_plan");
    return cs.done(create_interpolation_plan_call(handle, plan, nx, ny, nz, gridpos, mode));
}

static ffdas_error_t interpolation_preprocess_call(ffdas_handle_t handle,
                                                        ffdas_interpolation_plan_t plan,
                                                        int64_t num_querypos,
                                                        const float *querypos) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    if (!plan || !querypos) return FFDAS_ERROR_NULL_POINTER;
    if (num_querypos <= 0) return FFDAS_ERROR_INVALID_DIMS;
    FFDAS_TRY(handle->check_device());
    // drop an earlier preprocessed set (the stream may still read it)
    if (plan->pre_idx || plan->pre_w) {
        FFDAS_TRY(hip_status(hipStreamSynchronize(handle->stream), FFDAS_ERROR_CUDA_SYNC_FAILED));
        if (plan->pre_idx) (void)hipFree(plan->pre_idx);
        if (plan->pre_w) (void)hipFree(plan->pre_w);
        plan->pre_idx = nullptr;
        plan->pre_w = nullptr;
    }
    plan->pre = false;
    plan->pre_n = 0;
    FFDAS_TRY(alloc_lookup(plan->mode, num_querypos, &plan->pre_idx, &plan->pre_w));
    FFDAS_TRY(locate(*handle, *plan, querypos, num_querypos, plan->pre_idx, plan->pre_w));
    plan->pre = true;
    plan->pre_n = num_querypos;
    return FFDAS_SUCCESS;
}

extern "C" ffdas_error_t ffdas_interpolation_preprocess(ffdas_handle_t handle,
                                                        ffdas_interpolation_plan_t plan,
                                                        int64_t num_querypos,
                                                        const float *querypos) {
    ffdas::CallScope cs(handle, "ffdas_interpolation_preprocess");
    return cs.done(interpolation_preprocess_call(handle, plan, num_querypos, querypos));
}

static ffdas_error_t interpolation_call(ffdas_handle_t handle, ffdas_interpolation_plan_t plan,
                                             int64_t num_querypos, const float *querypos,
                                             ffdas_tensor_desc_t x_desc, const void *x, void *out,
                                             const void *fill_value) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    if (!plan || !x_desc || !x || !out || !fill_value) return FFDAS_ERROR_NULL_POINTER;
    const bool use_pre = plan->pre && plan->pre_n == num_querypos;
    if (!querypos && !use_pre) return FFDAS_ERROR_INVALID_ARGUMENT;
    if (num_querypos <= 0) return FFDAS_ERROR_INVALID_DIMS;
    const ffdas_tensor_desc &xd = *x_desc;
    const size_t nd = xd.dims.size();
    if (nd != 3 && nd != 4) return FFDAS_ERROR_INVALID_DIMS;
    const size_t o = nd - 3;
    if (xd.dims[o] != plan->nz || xd.dims[o + 1] != plan->ny || xd.dims[o + 2] != plan->nx)
        return FFDAS_ERROR_INVALID_DIMS;
    // packed layout: the kernels index x as [frame][vertex]
    int64_t expect = 1;
    for (size_t k = nd; k-- > 0;) {
        if (xd.dims[k] > 1 && xd.strides[k] != expect) return FFDAS_ERROR_NON_CONTIGUOUS;
        expect *= xd.dims[k];
    }
    const int64_t B = nd == 4 ? xd.dims[0] : 1;
    if (B <= 0) return FFDAS_ERROR_INVALID_DIMS;
    const int64_t bs = plan->nz * plan->ny * plan->nx;
    if ((B + kFramesPerThread - 1) / kFramesPerThread > 65535) return FFDAS_ERROR_DIMS_TOO_LARGE;
    FFDAS_TRY(handle->check_device());

    const int *idx = plan->pre_idx;
    const float *w = plan->pre_w;
    int *tidx = nullptr;
    float *tw = nullptr;
    if (!use_pre) {
        FFDAS_TRY(alloc_lookup(plan->mode, num_querypos, &tidx, &tw));
        const ffdas_error_t e = locate(*handle, *plan, querypos, num_querypos, tidx, tw);
        if (e != FFDAS_SUCCESS) {
            (void)hipFree(tidx);
            if (tw) (void)hipFree(tw);
            return e;
        }
        idx = tidx;
        w = tw;
    }
    ffdas_error_t err;
    switch (xd.dtype) {
        case FFDAS_R_32F: err = evaluate<float, 1>(*handle, plan->mode, x, out, idx, w, num_querypos, B, bs, fill_value); break;
        case FFDAS_C_32F: err = evaluate<float, 2>(*handle, plan->mode, x, out, idx, w, num_querypos, B, bs, fill_value); break;
        case FFDAS_R_64F: err = evaluate<double, 1>(*handle, plan->mode, x, out, idx, w, num_querypos, B, bs, fill_value); break;
        case FFDAS_C_64F: err = evaluate<double, 2>(*handle, plan->mode, x, out, idx, w, num_querypos, B, bs, fill_value); break;
        default: err = FFDAS_ERROR_UNSUPPORTED_TYPE;
    }
    if (tidx) {
        // the one-shot lookup is freed once the evaluation that reads it has run
        const hipError_t e = hipStreamSynchronize(handle->stream);
        (void)hipFree(tidx);
        if (tw) (void)hipFree(tw);
        if (err == FFDAS_SUCCESS) err = hip_status(e, FFDAS_ERROR_CUDA_SYNC_FAILED);
    }
    return err;
}

extern "C" ffdas_error_t ffdas_interpolation(ffdas_handle_t handle, ffdas_interpolation_plan_t plan,
                                             int64_t num_querypos, const float *querypos,
                                             ffdas_tensor_desc_t x_desc, const void *x, void *out,
                                             const void *fill_value) {
    ffdas::CallScope cs(handle, "ffdas_interpolation");
    return cs.done(interpolation_call(handle, plan, num_querypos, querypos, x_desc, x, out,
                                      fill_value));
}

extern "C" ffdas_error_t ffdas_destroy_interpolation_plan(ffdas_handle_t handle,
                                                          ffdas_interpolation_plan_t plan) {
    if (!plan) return FFDAS_SUCCESS;
    if (handle) (void)hipStreamSynchronize(handle->stream);
    free_plan_device(plan);
    delete plan;
    return FFDAS_SUCCESS;
}
