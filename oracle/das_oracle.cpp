// CPU oracle for the delay-and-sum hot path (`das`, `das_sparse`) and its adjoint.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
// compared against; it is never linked into, loaded by, or called from the
// product library (libffdas.so) or the Python package.  Only tests/, the
// smoke() check and bench.py's CPU-baseline leg use it.
//
// It restates the arithmetic definition of the reference's ALG1 kernel
// (src/das/das_alg1_kernels.cuh:41-106) and the documented semantics
// (docs/api/das.md:5,37-48,90-100; docs/conventions.md:9-35):
//
//   for every frame b and voxel n:
//     out[b,n] = sum_{j < K} sum_{m < C}  g(n,m,j) * lerp(x[b,m,s_j,k], x[b,m,s_j,k+1], t)
//   with
//     s_j   = j                 (dense das)          das_alg1_kernels.cuh:49
//           = sparse_idx[j,n]   (das_sparse)         das_alg1_kernels.cuh:51-53
//     d     = dstpos[n] - srcpos[m]                  das_alg1_kernels.cuh:59-62
//     r     = sqrt(max(|d|^2, 1e-7f))                das_alg1_kernels.cuh:64-66
//     tau   = scale * (offsets[j,n] + r)             (scale == 1 in the reference; the
//                                                     sound-speed scale is this build's
//                                                     additive extension)
//     k     = floor(tau), t = tau - k                das_device_utils.cuh:9-13
//     in    = 0 <= k < S-1                           das_alg1_kernels.cuh:70
//           && (d . srcdir[m].xyz) / r > srcdir[m].w (only if srcdir given, strict; :72-76)
//     g     = in ? weights[j,n] * (cos(wk*tau), sin(wk*tau)) : 0   (complex data, :81-85)
//           = in ? weights[j,n] : 0                                (real data, :88)
//     lerp  = x0 + t*(x1 - x0)                       core/math/vector.cuh:14-16
//     out   = beta*out + acc  if beta != 0 else acc  das_device_utils.cuh:31-39
//
// Deliberate departures from the reference's quirks (documented definition wins):
//   * offsets/weights/sparse_idx are indexed [j*N + n] for batched AND unbatched x
//     (the reference uses out.strides[0] = 0 for unbatched x, das_utils.h:64);
//   * x[k], x[k+1] are only read when 0 <= k < S-1 (no out-of-bounds read);
//   * a sparse index outside [0, T) contributes zero (the reference reads out of bounds).
//
// The adjoint (this build's extension) differentiates the same expression:
//   d out/d x[b,m,s,k]   = g*(1-t),   d out/d x[b,m,s,k+1] = g*t
//   d out/d tau          = w*((x1-x0) + i*wk*lerp) * e^{i wk tau}  (right derivative in tau)
//   d tau/d offsets      = scale,     d tau/d scale = offsets + r
//   d out/d weights      = e^{i wk tau} * lerp
// For complex data the gradients follow the real-loss convention g = dL/dRe + i dL/dIm:
//   grad_x = conj(coef) * g ;  grad_param = Re(conj(g) * d out/d param).
//
// Templated on the arithmetic type R: double is the correctness oracle, float is the
// CPU baseline ("the oracle compiled in FP32").
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr float kRadiusSqMin = 1e-7f;  // das_alg1_kernels.cuh:64 (rounded to float there)

struct Geometry {
    int B, C, T, S, N, K;
    int is_complex;
    const float *srcpos;   // (C,3)
    const float *srcdir;   // (C,4) or null
    float wavenum;
    double scale;
    const float *dstpos;   // (N,3)
    const float *offsets;  // (K,N)
    const float *weights;  // (K,N)
    const int *sparse;     // (K,N) or null (dense: K == T)
};

template <class R>
struct Tap {
    bool in;
    int64_t k;
    R t;
    R tau;
    R r;
    int s;
};

template <class R>
inline Tap<R> make_tap(const Geometry &g, int n, int m, int j) {
    Tap<R> tp;
    const float *y = g.dstpos + 3 * (int64_t)n;
    const float *p = g.srcpos + 3 * (int64_t)m;
    R dx = R(y[0]) - R(p[0]);
    R dy = R(y[1]) - R(p[1]);
    R dz = R(y[2]) - R(p[2]);
    R r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < R(kRadiusSqMin)) r2 = R(kRadiusSqMin);
    R r = std::sqrt(r2);
    R ofs = R(g.offsets[(int64_t)j * g.N + n]);
    R tau = R(g.scale) * (ofs + r);
    R kf = std::floor(tau);
    tp.tau = tau;
    tp.r = r;
    tp.t = tau - kf;
    tp.s = g.sparse ? g.sparse[(int64_t)j * g.N + n] : j;
    bool in = (kf >= R(0)) && (kf < R(g.S - 1)) && tp.s >= 0 && tp.s < g.T;
    tp.k = in ? (int64_t)kf : 0;
    if (in && g.srcdir) {
        const float *dv = g.srcdir + 4 * (int64_t)m;
        R proj = (dx * R(dv[0]) + dy * R(dv[1]) + dz * R(dv[2])) / r;
        in = proj > R(dv[3]);
    }
    tp.in = in;
    return tp;
}

template <class Fn>
void parallel_for(int64_t n, int nthreads, Fn fn) {
    if (nthreads <= 1 || n < 2) {
        fn(0, n, 0);
        return;
    }
    int nt = (int)std::min<int64_t>(nthreads, n);
    std::vector<std::thread> th;
    std::atomic<int64_t> next{0};
    const int64_t chunk = std::max<int64_t>(1, n / (nt * 16));
    for (int t = 0; t < nt; t++) {
        th.emplace_back([&, t] {
            for (;;) {
                int64_t lo = next.fetch_add(chunk);
                if (lo >= n) break;
                fn(lo, std::min(n, lo + chunk), t);
            }
        });
    }
    for (auto &x : th) x.join();
}

inline int64_t xrow(const Geometry &g, int b, int m, int s) {
    return (((int64_t)b * g.C + m) * g.T + s) * g.S;
}

// Forward: out (B,N)[,2]; beta applied as out = beta*out + acc (beta == 0: overwrite).
template <class R>
void das_forward(const Geometry &g, const R *x, R *out, double beta_re, double beta_im, int nthreads) {
    const int cw = g.is_complex ? 2 : 1;
    parallel_for(g.N, nthreads, [&](int64_t lo, int64_t hi, int) {
        std::vector<R> acc((size_t)g.B * cw);
        for (int64_t n = lo; n < hi; n++) {
            std::fill(acc.begin(), acc.end(), R(0));
            for (int j = 0; j < g.K; j++) {
                const R w = R(g.weights[(int64_t)j * g.N + n]);
                for (int m = 0; m < g.C; m++) {
                    Tap<R> tp = make_tap<R>(g, (int)n, m, j);
                    if (!tp.in) continue;
                    if (!g.is_complex) {
                        for (int b = 0; b < g.B; b++) {
                            const R *row = x + xrow(g, b, m, tp.s);
                            R x0 = row[tp.k], x1 = row[tp.k + 1];
                            acc[b] += w * (x0 + tp.t * (x1 - x0));
                        }
                    } else {
                        R ph = R(g.wavenum) * tp.tau;
                        R wr = w * std::cos(ph), wi = w * std::sin(ph);
                        for (int b = 0; b < g.B; b++) {
                            const R *row = x + 2 * xrow(g, b, m, tp.s);
                            R x0r = row[2 * tp.k], x0i = row[2 * tp.k + 1];
                            R x1r = row[2 * tp.k + 2], x1i = row[2 * tp.k + 3];
                            R lr = x0r + tp.t * (x1r - x0r);
                            R li = x0i + tp.t * (x1i - x0i);
                            acc[2 * b] += wr * lr - wi * li;
                            acc[2 * b + 1] += wr * li + wi * lr;
                        }
                    }
                }
            }
            for (int b = 0; b < g.B; b++) {
                R *o = out + ((int64_t)b * g.N + n) * cw;
                if (!g.is_complex) {
                    o[0] = (beta_re == 0.0) ? acc[b] : R(beta_re) * o[0] + acc[b];
                } else {
                    if (beta_re == 0.0 && beta_im == 0.0) {
                        o[0] = acc[2 * b];
                        o[1] = acc[2 * b + 1];
                    } else {
                        R orr = o[0], oi = o[1];
                        o[0] = R(beta_re) * orr - R(beta_im) * oi + acc[2 * b];
                        o[1] = R(beta_re) * oi + R(beta_im) * orr + acc[2 * b + 1];
                    }
                }
            }
        }
    });
}

// Adjoint.  gout (B,N)[,2].  Outputs may be null.  gx (B,C,T,S)[,2] is overwritten,
// goffsets/gweights (K,N) overwritten, gscale[0] overwritten.
template <class R>
void das_adjoint(const Geometry &g, const R *x, const R *gout, R *gx, R *goffsets, R *gweights,
                 R *gscale, int nthreads) {
    const int cw = g.is_complex ? 2 : 1;
    if (gx) {
        std::memset(gx, 0, sizeof(R) * (size_t)g.B * g.C * g.T * g.S * cw);
        // parallel over channels: every channel owns its own rows of gx
        parallel_for(g.C, nthreads, [&](int64_t lo, int64_t hi, int) {
            for (int64_t m = lo; m < hi; m++) {
                for (int64_t n = 0; n < g.N; n++) {
                    for (int j = 0; j < g.K; j++) {
                        Tap<R> tp = make_tap<R>(g, (int)n, (int)m, j);
                        if (!tp.in) continue;
                        const R w = R(g.weights[(int64_t)j * g.N + n]);
                        if (!g.is_complex) {
                            for (int b = 0; b < g.B; b++) {
                                R go = gout[(int64_t)b * g.N + n];
                                R *row = gx + xrow(g, b, (int)m, tp.s);
                                row[tp.k] += go * w * (R(1) - tp.t);
                                row[tp.k + 1] += go * w * tp.t;
                            }
                        } else {
                            R ph = R(g.wavenum) * tp.tau;
                            R cr = w * std::cos(ph), ci = w * std::sin(ph);
                            for (int b = 0; b < g.B; b++) {
                                R gr = gout[2 * ((int64_t)b * g.N + n)];
                                R gi = gout[2 * ((int64_t)b * g.N + n) + 1];
                                // conj(c) * g
                                R vr = cr * gr + ci * gi;
                                R vi = cr * gi - ci * gr;
                                R *row = gx + 2 * xrow(g, b, (int)m, tp.s);
                                row[2 * tp.k] += vr * (R(1) - tp.t);
                                row[2 * tp.k + 1] += vi * (R(1) - tp.t);
                                row[2 * tp.k + 2] += vr * tp.t;
                                row[2 * tp.k + 3] += vi * tp.t;
                            }
                        }
                    }
                }
            }
        });
    }
    if (goffsets || gweights || gscale) {
        int nt = std::max(1, nthreads);
        std::vector<R> part_scale(nt, R(0));
        parallel_for(g.N, nthreads, [&](int64_t lo, int64_t hi, int tid) {
            R ps = 0;
            for (int64_t n = lo; n < hi; n++) {
                for (int j = 0; j < g.K; j++) {
                    const R w = R(g.weights[(int64_t)j * g.N + n]);
                    R dtau = 0, dw = 0, dscale = 0;
                    for (int m = 0; m < g.C; m++) {
                        Tap<R> tp = make_tap<R>(g, (int)n, m, j);
                        if (!tp.in) continue;
                        R dt_m = 0;  // sum_b Re(conj(g) * d out/d tau) / w
                        R dw_m = 0;
                        if (!g.is_complex) {
                            for (int b = 0; b < g.B; b++) {
                                const R *row = x + xrow(g, b, m, tp.s);
                                R x0 = row[tp.k], x1 = row[tp.k + 1];
                                R go = gout[(int64_t)b * g.N + n];
                                dt_m += go * (x1 - x0);
                                dw_m += go * (x0 + tp.t * (x1 - x0));
                            }
                        } else {
                            R ph = R(g.wavenum) * tp.tau;
                            R er = std::cos(ph), ei = std::sin(ph);
                            R wk = R(g.wavenum);
                            for (int b = 0; b < g.B; b++) {
                                const R *row = x + 2 * xrow(g, b, m, tp.s);
                                R x0r = row[2 * tp.k], x0i = row[2 * tp.k + 1];
                                R x1r = row[2 * tp.k + 2], x1i = row[2 * tp.k + 3];
                                R lr = x0r + tp.t * (x1r - x0r);
                                R li = x0i + tp.t * (x1i - x0i);
                                // e * lerp
                                R pr = er * lr - ei * li, pi = er * li + ei * lr;
                                // e * ((x1-x0) + i*wk*lerp)
                                R qr0 = (x1r - x0r) - wk * li, qi0 = (x1i - x0i) + wk * lr;
                                R qr = er * qr0 - ei * qi0, qi = er * qi0 + ei * qr0;
                                R gr = gout[2 * ((int64_t)b * g.N + n)];
                                R gi = gout[2 * ((int64_t)b * g.N + n) + 1];
                                dt_m += gr * qr + gi * qi;
                                dw_m += gr * pr + gi * pi;
                            }
                        }
                        dtau += w * dt_m;
                        dw += dw_m;
                        dscale += w * dt_m * (R(g.offsets[(int64_t)j * g.N + n]) + tp.r);
                    }
                    if (goffsets) goffsets[(int64_t)j * g.N + n] = R(g.scale) * dtau;
                    if (gweights) gweights[(int64_t)j * g.N + n] = dw;
                    ps += dscale;
                }
            }
            part_scale[tid] += ps;
        });
        if (gscale) {
            R s = 0;
            for (R v : part_scale) s += v;
            gscale[0] = s;
        }
    }
}

Geometry make_geometry(int B, int C, int T, int S, int N, int K, int is_complex, const float *srcpos,
                       const float *srcdir, float wavenum, double scale, const float *dstpos,
                       const float *offsets, const float *weights, const int *sparse) {
    Geometry g;
    g.B = B; g.C = C; g.T = T; g.S = S; g.N = N; g.K = K;
    g.is_complex = is_complex;
    g.srcpos = srcpos; g.srcdir = srcdir; g.wavenum = wavenum; g.scale = scale;
    g.dstpos = dstpos; g.offsets = offsets; g.weights = weights; g.sparse = sparse;
    return g;
}

int check_args(int B, int C, int T, int S, int N, int K, const float *sp, const float *dp,
               const float *of, const float *we) {
    if (B <= 0 || C <= 0 || T <= 0 || S <= 0 || N <= 0 || K <= 0) return 1;
    if (!sp || !dp || !of || !we) return 2;
    return 0;
}

}  // namespace

extern "C" {

// real_bits: 64 -> x/out/grads are double arrays, arithmetic in double (the oracle);
//            32 -> float arrays, arithmetic in float (the CPU baseline).
// sparse == null means dense das with K == T.
int ffdas_oracle_das(int real_bits, int nthreads, int B, int C, int T, int S, int N, int K,
                     int is_complex, const float *srcpos, const float *srcdir, float wavenum,
                     double scale, const void *x, const float *dstpos, const float *offsets,
                     const float *weights, const int *sparse, double beta_re, double beta_im,
                     void *out) {
    int e = check_args(B, C, T, S, N, K, srcpos, dstpos, offsets, weights);
    if (e) return e;
    if (!sparse && K != T) return 3;
    Geometry g = make_geometry(B, C, T, S, N, K, is_complex, srcpos, srcdir, wavenum, scale, dstpos,
                               offsets, weights, sparse);
    if (real_bits == 64)
        das_forward<double>(g, (const double *)x, (double *)out, beta_re, beta_im, nthreads);
    else if (real_bits == 32)
        das_forward<float>(g, (const float *)x, (float *)out, beta_re, beta_im, nthreads);
    else
        return 4;
    return 0;
}

int ffdas_oracle_das_adjoint(int real_bits, int nthreads, int B, int C, int T, int S, int N, int K,
                             int is_complex, const float *srcpos, const float *srcdir,
                             float wavenum, double scale, const void *x, const float *dstpos,
                             const float *offsets, const float *weights, const int *sparse,
                             const void *gout, void *gx, void *goffsets, void *gweights,
                             void *gscale) {
    int e = check_args(B, C, T, S, N, K, srcpos, dstpos, offsets, weights);
    if (e) return e;
    if (!sparse && K != T) return 3;
    Geometry g = make_geometry(B, C, T, S, N, K, is_complex, srcpos, srcdir, wavenum, scale, dstpos,
                               offsets, weights, sparse);
    if (real_bits == 64)
        das_adjoint<double>(g, (const double *)x, (const double *)gout, (double *)gx,
                            (double *)goffsets, (double *)gweights, (double *)gscale, nthreads);
    else if (real_bits == 32)
        das_adjoint<float>(g, (const float *)x, (const float *)gout, (float *)gx, (float *)goffsets,
                           (float *)gweights, (float *)gscale, nthreads);
    else
        return 4;
    return 0;
}

int ffdas_oracle_hardware_threads(void) { return (int)std::thread::hardware_concurrency(); }

}  // extern "C"
