// ffdas_truncate_rank: rank truncation filter (reference: src/truncate_rank/truncate_rank.cu:6-42,
// src/truncate_rank/truncate_rank_impl.cuh:382-531; Python bindings/python/ffdas/
// truncate_rank.py:28-82).  Gram matrix and projection run on the kernels of
// rank_kernels.hip; the p x p Hermitian eigendecomposition runs on the one-workgroup Jacobi
// kernel for p <= 64; above, FP32 Gram matrices up to p = 224 (real) or 160 (complex) take the
// tridiagonal path (one-workgroup Householder reduction, the chosen eigenpairs by bisection
// and inverse iteration, their back-transformation), and the rest go to rocSOLVER's
// syevd/heevd, loaded on first use so that the library does not pull rocSOLVER in at load time.
#include <dlfcn.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "internal.h"

namespace ffdas {
namespace {

struct Solver {
    bool ok = false;
    decltype(&rocblas_create_handle) create = nullptr;
    decltype(&rocblas_destroy_handle) destroy = nullptr;
    decltype(&rocblas_set_stream) set_stream = nullptr;
    decltype(&rocsolver_ssyevd) ssyevd = nullptr;
    decltype(&rocsolver_dsyevd) dsyevd = nullptr;
    decltype(&rocsolver_cheevd) cheevd = nullptr;
    decltype(&rocsolver_zheevd) zheevd = nullptr;
};

void *open_lib(const char *name) {
    const char *root = std::getenv("ROCM_PATH");
    const std::string dir = std::string(root && *root ? root : "/opt/rocm") + "/lib/";
    void *h = dlopen((dir + name).c_str(), RTLD_NOW | RTLD_LOCAL);
    return h ? h : dlopen(name, RTLD_NOW | RTLD_LOCAL);
}

const Solver &solver() {
    static Solver s;
    static std::once_flag once;
    std::call_once(once, [] {
        void *blas = open_lib("librocblas.so.5");
        void *sol = open_lib("librocsolver.so.0");
        if (!blas || !sol) return;
        s.create = (decltype(s.create))dlsym(blas, "rocblas_create_handle");
        s.destroy = (decltype(s.destroy))dlsym(blas, "rocblas_destroy_handle");
        s.set_stream = (decltype(s.set_stream))dlsym(blas, "rocblas_set_stream");
        s.ssyevd = (decltype(s.ssyevd))dlsym(sol, "rocsolver_ssyevd");
        s.dsyevd = (decltype(s.dsyevd))dlsym(sol, "rocsolver_dsyevd");
        s.cheevd = (decltype(s.cheevd))dlsym(sol, "rocsolver_cheevd");
        s.zheevd = (decltype(s.zheevd))dlsym(sol, "rocsolver_zheevd");
        s.ok = s.create && s.destroy && s.set_stream && s.ssyevd && s.dsyevd && s.cheevd &&
               s.zheevd;
    });
    return s;
}

size_t align_up(size_t b) { return (b + 255) & ~size_t(255); }

constexpr int kJacobiSweeps = 30;
constexpr int kJacobiMaxP = 64;
constexpr int kJacobiOverTridiagP = 16;   // FP32 p at or below which Jacobi stays ahead

// The solver's status word, read on the host: the one synchronisation of the call, as the
// reference's (truncate_rank_impl.cuh:82-87).
ffdas_error_t check_info(ffdas_context &h, const int *info, ffdas_error_t fail) {
    if (!h.status_host && hipHostMalloc((void **)&h.status_host, sizeof(int)) != hipSuccess) {
        h.status_host = nullptr;
        return FFDAS_ERROR_ALLOCATION_FAILED;
    }
    *h.status_host = -1;
    if (hipMemcpyAsync(h.status_host, info, sizeof(int), hipMemcpyDeviceToHost, h.stream) !=
        hipSuccess)
        return FFDAS_ERROR_CUDA_MEMCPY_FAILED;
    if (hipStreamSynchronize(h.stream) != hipSuccess) return FFDAS_ERROR_CUDA_SYNC_FAILED;
    return *h.status_host == 0 ? FFDAS_SUCCESS : fail;
}

enum class EigPath { jacobi, tridiag, rocsolver };

// FP32 takes the tridiagonal path wherever its tile fits (real p <= 224, complex p <= 160)
// from p = 17 on: measured on gfx950 (whole calls), 64 x 64^3 complex 0.44 ms against the
// Jacobi kernel's 0.70 ms and heevd's 1.45 ms, real 0.29 / 0.52 / 1.27 ms; p = 24 real 0.131
// against Jacobi's 0.165 ms; p = 120 complex 1.15 ms against heevd's 2.62 ms; 200 x 96^3 real
// 1.79 ms against syevd's 4.43 ms; at p = 8 Jacobi is ahead (0.107 against 0.120 ms).  FP64
// (and FP32 up to p = 16) takes the one-workgroup Jacobi kernel (A and V in LDS) up to p = 64
// (at p = 64 it beat syevd/heevd: 0.74 / 1.19 ms real, 1.04 / 1.30 ms complex in FP32; its p^3
// cost per sweep loses above), and rocSOLVER's syevd/heevd beyond.
// FFDAS_RANK_EIG=jacobi (any p whose A and V fit in LDS), =tridiag (FP32 within those p) or
// =rocsolver overrides; a forced path that does not fit falls back to rocSOLVER's syevd/heevd.
EigPath eig_path(const RankPlan &pl, int lds_bytes) {
    const char *force = std::getenv("FFDAS_RANK_EIG");
    const std::string want = force ? force : "";
    const bool tri = rank_tridiag_fits(pl, lds_bytes);
    if (want == "jacobi")
        return rank_jacobi_fits(pl, lds_bytes) ? EigPath::jacobi : EigPath::rocsolver;
    if (want == "tridiag") return tri ? EigPath::tridiag : EigPath::rocsolver;
    if (want == "rocsolver") return EigPath::rocsolver;
    const bool jac = pl.p <= kJacobiMaxP && rank_jacobi_fits(pl, lds_bytes);
    if (tri && !(jac && pl.p <= kJacobiOverTridiagP)) return EigPath::tridiag;
    if (jac) return EigPath::jacobi;
    return EigPath::rocsolver;
}

// Device scratch of the tridiagonal path: reflectors U (p x p elements, packed into the first
// p (p - 1) / 2), the tridiagonal's eigenvectors Y (real p x r), tau (p elements) and 8 status
// words.
size_t tridiag_bytes(const RankPlan &pl) {
    const size_t p = (size_t)pl.p, c = pl.cplx ? 2 : 1;
    return align_up(c * p * p * 4) + align_up(p * (size_t)pl.r * 4) + align_up(c * p * 4) +
           align_up(8 * sizeof(int));
}

ffdas_error_t blas_handle(ffdas_context &h, const Solver &s, rocblas_handle *out) {
    if (!h.blas) {
        rocblas_handle bh = nullptr;
        if (s.create(&bh) != rocblas_status_success) return FFDAS_ERROR_CUSOLVER;
        h.blas = bh;
    }
    *out = (rocblas_handle)h.blas;
    if (s.set_stream(*out, h.stream) != rocblas_status_success) return FFDAS_ERROR_CUSOLVER;
    return FFDAS_SUCCESS;
}

// The tridiagonal path: T = Q^H H Q (rank_tridiag_kernel), the eigenpairs of T for the columns
// the projection reads (rank_tridiag_eigvec_kernel), then z = Q y into those columns of H
// (rank_backtransform_kernel, which also sets *info).  The other columns of H keep A.
ffdas_error_t eigh_tridiag(ffdas_context &h, const RankPlan &pl, void *H, void *D, void *E,
                           int *info, void *scratch) {
    const size_t p = (size_t)pl.p;
    char *w = (char *)scratch;
    const size_t c = pl.cplx ? 2 : 1;
    float *U = (float *)w;
    float *Y = (float *)(w + align_up(c * p * p * 4));
    float *tau = (float *)(w + align_up(c * p * p * 4) + align_up(p * (size_t)pl.r * 4));
    int *stat = (int *)((char *)tau + align_up(c * p * 4));
    FFDAS_TRY(launch_rank_tridiag(h.stream, pl, h.lds_bytes, H, D, E, U, tau, stat));
    FFDAS_TRY(launch_rank_tridiag_eigvec(h.stream, pl, D, E, Y, stat));
    FFDAS_TRY(launch_rank_backtransform(h.stream, pl, h.lds_bytes, U, tau, Y, stat, H, info));
    if (std::getenv("FFDAS_RANK_DEBUG")) {
        int hs = 0;
        if (hipMemcpyAsync(&hs, stat, sizeof(int), hipMemcpyDeviceToHost, h.stream) == hipSuccess &&
            hipStreamSynchronize(h.stream) == hipSuccess)
            std::fprintf(stderr, "ffdas: tridiagonal p=%d r=%d status %d\n", pl.p, pl.r, hs);
    }
    return FFDAS_SUCCESS;
}

// Eigendecomposition of the column-major p x p Hermitian H in place (eigenvectors in the
// columns, eigenvalues ascending; the tridiagonal path forms only the columns the projection
// reads), queued on the stream; *fail is the error to report if the solver's status word
// (read by the caller once the call's work is done) is not zero.
ffdas_error_t eigh(ffdas_context &h, const RankPlan &pl, EigPath path, void *H, void *D, void *E,
                   int *info, void *scratch, ffdas_error_t *fail) {
    if (path == EigPath::tridiag) {
        FFDAS_TRY(eigh_tridiag(h, pl, H, D, E, info, scratch));
        *fail = FFDAS_ERROR_CONVERGENCE_FAILED;
        return FFDAS_SUCCESS;
    }
    if (path == EigPath::jacobi) {
        int sweeps = kJacobiSweeps;
        if (const char *e = std::getenv("FFDAS_RANK_SWEEPS")) sweeps = std::max(1, std::atoi(e));
        FFDAS_TRY(launch_rank_jacobi(h.stream, pl, H, info, sweeps));
        if (std::getenv("FFDAS_RANK_DEBUG")) {
            int hi[2] = {0, 0};
            if (hipMemcpyAsync(hi, info, sizeof(hi), hipMemcpyDeviceToHost, h.stream) == hipSuccess &&
                hipStreamSynchronize(h.stream) == hipSuccess)
                std::fprintf(stderr, "ffdas: Jacobi p=%d sweeps=%d info=%d\n", pl.p, hi[1], hi[0]);
        }
        *fail = FFDAS_ERROR_CONVERGENCE_FAILED;
        return FFDAS_SUCCESS;
    }
    const Solver &s = solver();
    if (!s.ok) return FFDAS_ERROR_CUSOLVER;
    rocblas_handle bh = nullptr;
    FFDAS_TRY(blas_handle(h, s, &bh));
    const rocblas_evect ev = rocblas_evect_original;
    const rocblas_fill up = rocblas_fill_upper;
    rocblas_status st;
    if (pl.dbl && pl.cplx)
        st = s.zheevd(bh, ev, up, pl.p, (rocblas_double_complex *)H, pl.p, (double *)D,
                      (double *)E, info);
    else if (pl.dbl)
        st = s.dsyevd(bh, ev, up, pl.p, (double *)H, pl.p, (double *)D, (double *)E, info);
    else if (pl.cplx)
        st = s.cheevd(bh, ev, up, pl.p, (rocblas_float_complex *)H, pl.p, (float *)D,
                      (float *)E, info);
    else
        st = s.ssyevd(bh, ev, up, pl.p, (float *)H, pl.p, (float *)D, (float *)E, info);
    if (st != rocblas_status_success) return FFDAS_ERROR_CUSOLVER;
    *fail = FFDAS_ERROR_CUSOLVER;
    return FFDAS_SUCCESS;
}

bool overlaps(const void *a, size_t na, const void *b, size_t nb) {
    const char *x = (const char *)a, *y = (const char *)b;
    return x < y + nb && y < x + na;
}

}  // namespace

void release_rank_solver(ffdas_context &h) {
    if (h.blas) {
        const Solver &s = solver();
        if (s.ok) s.destroy((rocblas_handle)h.blas);
        h.blas = nullptr;
    }
}

}  // namespace ffdas

using namespace ffdas;

static ffdas_error_t truncate_rank_call(ffdas_handle_t handle, ffdas_tensor_desc_t x_desc,
                                             const void *x, int64_t start, int64_t stop,
                                             ffdas_tensor_desc_t out_desc, void *out) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    if (!x_desc || !out_desc) return FFDAS_ERROR_NULL_POINTER;
    FFDAS_TRY(handle->check_device());
    const ffdas_tensor_desc xd = *x_desc, od = *out_desc;
    if (xd.dtype != od.dtype) return FFDAS_ERROR_UNSUPPORTED_TYPE;
    RankPlan pl;
    switch (xd.dtype) {
        case FFDAS_R_32F: break;
        case FFDAS_C_32F: pl.cplx = true; break;
        case FFDAS_R_64F: pl.dbl = true; break;
        case FFDAS_C_64F: pl.dbl = pl.cplx = true; break;
        default: return FFDAS_ERROR_UNSUPPORTED_TYPE;
    }
    if (xd.dims.size() != 2 || od.dims != xd.dims) return FFDAS_ERROR_INVALID_DIMS;
    if (xd.dims[0] > INT_MAX || xd.dims[1] > INT_MAX) return FFDAS_ERROR_INVALID_DIMS;
    const int64_t m = xd.dims[0], n = xd.dims[1];
    const int64_t p = m < n ? m : n;
    // the reference validates the range in its bindings (truncate_rank.py:58-66); the C entry
    // point checks it too, since the kernels index eigenvector columns with it
    if (p < 1 || start < 0 || start >= p || stop <= start || stop > p)
        return FFDAS_ERROR_INVALID_ARGUMENT;
    if (!x || !out) return FFDAS_ERROR_NULL_POINTER;
    const int64_t xs0 = xd.strides[0], xs1 = xd.strides[1];
    const int64_t os0 = od.strides[0], os1 = od.strides[1];
    if (xs0 < 0 || xs1 < 0) return FFDAS_ERROR_NON_CONTIGUOUS;
    // out must not overlap itself: row-major or column-major dense layouts, or any strides
    // that keep distinct elements apart
    const bool out_ok = (os1 >= 1 && os0 >= os1 * n) || (os0 >= 1 && os1 >= os0 * m) || m * n <= 1;
    if (!out_ok) return FFDAS_ERROR_NON_CONTIGUOUS;

    const size_t esz = (pl.dbl ? 8 : 4) * (pl.cplx ? 2 : 1);
    const size_t rsz = pl.dbl ? 8 : 4;
    hipStream_t st = handle->stream;
    const int64_t kept = stop - start, dropped = p - kept;
    const size_t xspan = (size_t)((m - 1) * xs0 + (n - 1) * xs1 + 1) * esz;
    const size_t ospan = (size_t)((m - 1) * os0 + (n - 1) * os1 + 1) * esz;
    const bool same_layout = x == out && xs0 == os0 && xs1 == os1;
    // out overlapping x in another layout: every kernel below would read x after writing it
    const bool alias = !same_layout && overlaps(x, xspan, out, ospan);
    if (dropped == 0) {
        // every component kept: the reconstruction is x itself
        if (same_layout) return FFDAS_SUCCESS;
        if (!alias) return launch_rank_copy(st, pl.dbl, pl.cplx, x, xs0, xs1, out, os0, os1, m, n);
        void *tmp = nullptr;
        FFDAS_TRY(handle->workspace((size_t)m * n * esz, &tmp));
        handle->xcopy.valid = false;
        FFDAS_TRY(launch_rank_copy(st, pl.dbl, pl.cplx, x, xs0, xs1, tmp, n, 1, m, n));
        return launch_rank_copy(st, pl.dbl, pl.cplx, tmp, n, 1, out, os0, os1, m, n);
    }

    pl.m = m;
    pl.n = n;
    pl.p = (int)p;
    const bool rows = m <= n;   // Gram of the rows (m x m), projection of the columns
    pl.L = rows ? n : m;
    pl.gsa = rows ? xs0 : xs1;
    pl.gsk = rows ? xs1 : xs0;
    pl.xsj = rows ? xs1 : xs0;
    pl.xsf = rows ? xs0 : xs1;
    pl.osj = rows ? os1 : os0;
    pl.osf = rows ? os0 : os1;
    const auto aligned = [&](const void *ptr, int64_t stride_between) {
        return ((uintptr_t)ptr % 16 == 0) && ((stride_between * (int64_t)esz) % 16 == 0);
    };
    pl.gvec = pl.gsk == 1 && aligned(x, pl.gsa);
    pl.keep = kept <= dropped;
    pl.r = (int)(pl.keep ? kept : dropped);
    // eigenvalues ascend: component i (descending) is column p-1-i
    if (pl.keep) {
        pl.q_split = pl.r;
        pl.c_first = (int)(p - stop);
        pl.c_second = 0;
    } else {
        pl.q_split = (int)(p - stop);   // columns [0, p-stop) and [p-start, p)
        pl.c_first = 0;
        pl.c_second = (int)(p - start);
    }
    rank_plan_launch_shape(pl, handle->cu_count);

    // x is read again by later passes, and by other lanes when out aliases it in another
    // layout: such an output needs a private copy of x (out == x in the same layout and one
    // pass is safe: each lane reads its elements before it writes them)
    const bool need_copy = alias || (same_layout && pl.passes > 1);

    const size_t b_part = align_up(pl.part_bytes);
    const size_t b_H = align_up((size_t)p * p * esz);
    const size_t b_D = align_up((size_t)p * rsz);
    const size_t b_info = align_up(sizeof(int));
    const size_t b_Q = align_up((size_t)pl.rp * pl.pp * esz);
    const size_t b_x = need_copy ? align_up((size_t)m * n * esz) : 0;
    const EigPath path = eig_path(pl, handle->lds_bytes);
    const size_t b_eig = path == EigPath::tridiag ? tridiag_bytes(pl) : 0;
    void *ws = nullptr;
    FFDAS_TRY(handle->workspace(b_part + b_H + 2 * b_D + b_info + 2 * b_Q + b_x + b_eig, &ws));
    handle->xcopy.valid = false;   // the converted-x record of the delay-and-sum calls
    char *w = (char *)ws;
    void *part = w;
    void *H = w + b_part;
    void *D = w + b_part + b_H;
    void *E = w + b_part + b_H + b_D;
    int *info = (int *)(w + b_part + b_H + 2 * b_D);
    void *Qh = w + b_part + b_H + 2 * b_D + b_info;
    void *Qn = (char *)Qh + b_Q;
    void *eig_scratch = (char *)Qn + b_Q + b_x;
    const void *xs = x;
    if (need_copy) {
        void *xc = (char *)Qn + b_Q;
        FFDAS_TRY(launch_rank_copy(st, pl.dbl, pl.cplx, x, xs0, xs1, xc, n, 1, m, n));
        xs = xc;
        pl.xsj = rows ? 1 : n;
        pl.xsf = rows ? n : 1;
    }
    pl.xvec = pl.xsf == 1 && aligned(xs, pl.xsj);
    pl.ovec = pl.osf == 1 && aligned(out, pl.osj);

    FFDAS_TRY(launch_rank_gram(st, pl, x, part, H));
    // the projection is queued behind the eigensolver; its status word is read once at the
    // end (a failed solve leaves out undefined and returns the solver's error)
    ffdas_error_t fail = FFDAS_SUCCESS;
    FFDAS_TRY(eigh(*handle, pl, path, H, D, E, info, eig_scratch, &fail));
    FFDAS_TRY(launch_rank_project(st, pl, H, xs, out, Qh, Qn));
    return check_info(*handle, info, fail);
}

extern "C" ffdas_error_t ffdas_truncate_rank(ffdas_handle_t handle, ffdas_tensor_desc_t x_desc,
                                             const void *x, int64_t start, int64_t stop,
                                             ffdas_tensor_desc_t out_desc, void *out) {
    ffdas::CallScope cs(handle, "ffdas_truncate_rank");
    return cs.done(truncate_rank_call(handle, x_desc, x, start, stop, out_desc, out));
}
