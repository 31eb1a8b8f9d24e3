// C-ABI entry points of the delay-and-sum path: ffdas_das, ffdas_das_sparse (reference
// src/das/das.cu:9-83, src/das/das_sparse.cu:6-80), the scaled/sparse forward ffdas_das_ex,
// the adjoint ffdas_das_backward and the batched forward ffdas_das_batched
// (ffdas_contiguous_copy is in copy.hip).
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace ffdas {

size_t storage_bytes(Storage s) {
    switch (s) {
        case Storage::F16: case Storage::BF16: return 2;
        case Storage::F32: return 4;
        case Storage::F64: return 8;
    }
    return 0;
}

namespace {

// beta is a host scalar of the output type (src/das/das_impl.cuh:89)
void read_beta(const void *beta, const DasProblem &p, double &re, double &im) {
    re = im = 0.0;
    if (!beta) return;
    if (p.acc64) {
        const double *b = static_cast<const double *>(beta);
        re = b[0];
        im = p.cplx ? b[1] : 0.0;
    } else {
        const float *b = static_cast<const float *>(beta);
        re = b[0];
        im = p.cplx ? b[1] : 0.0;
    }
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

size_t cast_bytes(const DasProblem &p) {
    if (p.fi_frames > 0) return align_up(fi_bytes(p));
    if (!p.x_needs_copy) return 0;
    const size_t cw = p.cplx ? 2 : 1;
    return align_up((size_t)p.B * p.C * p.T * p.S * cw * storage_bytes(p.k_storage));
}

// Produce the view of x the kernels gather from: the caller's tensor when it already has
// the kernel storage type, otherwise a contiguous converted copy in workspace `ws`
// (cast_bytes(p) bytes).
// Whether the copy recorded in h.xcopy was made from this x for this problem.
bool xcopy_matches(const ffdas_context &h, const ffdas_tensor_desc &xd, const DasProblem &p,
                   const void *x, void *ws) {
    const auto &c = h.xcopy;
    return c.valid && c.x == x && c.ws == ws && c.dims == xd.dims && c.strides == xd.strides &&
           c.dtype == (int)xd.dtype && c.storage == (int)p.k_storage &&
           c.fi_frames == p.fi_frames && c.cast == (p.x_needs_copy ? 1 : 0) &&
           c.sparse == (p.sparse ? 1 : 0);
}

void record_xcopy(ffdas_context &h, const ffdas_tensor_desc &xd, const DasProblem &p,
                  const void *x, void *ws) {
    auto &c = h.xcopy;
    c.valid = p.fi_frames > 0 || p.x_needs_copy;
    c.x = x;
    c.ws = ws;
    c.dims = xd.dims;
    c.strides = xd.strides;
    c.dtype = (int)xd.dtype;
    c.storage = (int)p.k_storage;
    c.fi_frames = p.fi_frames;
    c.cast = p.x_needs_copy ? 1 : 0;
    c.sparse = p.sparse ? 1 : 0;     // the frame-inner layout depends on it
}

// The view of an already converted copy at `ws` (no launch).
XView copy_view(const DasProblem &p, const void *ws) {
    XView xv;
    xv.ptr = ws;
    if (p.fi_frames <= 0) {
        xv.st = p.S;
        xv.sc = (int64_t)p.T * p.S;
        xv.sb = (int64_t)p.C * p.T * p.S;
    }
    return xv;
}

ffdas_error_t prepare_x(ffdas_context &h, const DasProblem &p, const void *x, void *ws,
                        XView &xv) {
    if (p.fi_frames > 0) {
        FFDAS_TRY(launch_relayout_fi(h, p, x, ws));
        xv.ptr = ws;
        return FFDAS_SUCCESS;
    }
    if (!p.x_needs_copy) {
        xv.ptr = x;
        xv.sb = p.xs_b;
        xv.sc = p.xs_c;
        xv.st = p.xs_t;
        return FFDAS_SUCCESS;
    }
    FFDAS_TRY(launch_cast(h, p, x, ws));
    xv.ptr = ws;
    xv.st = p.S;
    xv.sc = (int64_t)p.T * p.S;
    xv.sb = (int64_t)p.C * p.T * p.S;
    return FFDAS_SUCCESS;
}

}  // namespace

ffdas_error_t das_forward_with(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                               float wavenum, float scale, ffdas_tensor_desc_t x_desc,
                               const void *x, const float *dstpos, const float *offsets,
                               const float *weights, bool sparse_mode, int sparse_count,
                               const int *sparse_indices, const void *beta,
                               ffdas_tensor_desc_t out_desc, void *out,
                               ffdas_compute_type_t compute_type, ffdas_alg_t alg,
                               const WorkspaceFn &get_ws, bool record_copy) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    FFDAS_TRY(handle->check_device());
    if (!x_desc || !out_desc) return FFDAS_ERROR_NULL_POINTER;
    const ffdas_tensor_desc xd = *x_desc, od = *out_desc;   // copied at entry
    DasProblem p;
    FFDAS_TRY(plan_das(xd, od, srcpos, dstpos, offsets, weights, sparse_mode, sparse_count,
                       sparse_indices, compute_type, alg, p));
    resolve_tune(*handle, p, srcdir != nullptr);
    p.fi_frames = fi_frames_for(p, alg);
    if (!x || !out) return FFDAS_ERROR_NULL_POINTER;
    if (!(scale > 0.0f) || !std::isfinite(scale)) return FFDAS_ERROR_INVALID_ARGUMENT;
    double br, bi;
    read_beta(beta, p, br, bi);
    void *ws = nullptr;
    FFDAS_TRY(get_ws(cast_bytes(p), &ws));
    XView xv;
    FFDAS_TRY(prepare_x(*handle, p, x, ws, xv));
    if (record_copy) record_xcopy(*handle, xd, p, x, ws);
    DasGeomArgs a;
    a.srcpos = srcpos; a.srcdir = srcdir; a.dstpos = dstpos; a.offsets = offsets;
    a.weights = weights; a.sparse_indices = sparse_indices; a.wavenum = wavenum; a.scale = scale;
    return launch_forward(*handle, p, a, xv, out, br, bi);
}

namespace {

ffdas_error_t das_forward_common(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                                 float wavenum, float scale, ffdas_tensor_desc_t x_desc,
                                 const void *x, const float *dstpos, const float *offsets,
                                 const float *weights, bool sparse_mode, int sparse_count,
                                 const int *sparse_indices, const void *beta,
                                 ffdas_tensor_desc_t out_desc, void *out,
                                 ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    const WorkspaceFn ws = [handle](size_t bytes, void **ptr) {
        return handle->workspace(bytes, ptr);
    };
    return das_forward_with(handle, srcpos, srcdir, wavenum, scale, x_desc, x, dstpos, offsets,
                            weights, sparse_mode, sparse_count, sparse_indices, beta, out_desc,
                            out, compute_type, alg, ws, true);
}

}  // namespace
}  // namespace ffdas

using namespace ffdas;

extern "C" {

static ffdas_error_t das_call(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                        float wavenum, ffdas_tensor_desc_t x_desc, const void *x,
                        const float *dstpos, const float *offsets, const float *weights,
                        const void *beta, ffdas_tensor_desc_t out_desc, void *out,
                        ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    return das_forward_common(handle, srcpos, srcdir, wavenum, 1.0f, x_desc, x, dstpos, offsets,
                              weights, false, 0, nullptr, beta, out_desc, out, compute_type, alg);
}

ffdas_error_t ffdas_das(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                        float wavenum, ffdas_tensor_desc_t x_desc, const void *x,
                        const float *dstpos, const float *offsets, const float *weights,
                        const void *beta, ffdas_tensor_desc_t out_desc, void *out,
                        ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    ffdas::CallScope cs(handle, "ffdas_das");
    return cs.done(das_call(handle, srcpos, srcdir, wavenum, x_desc, x, dstpos, offsets, weights,
                            beta, out_desc, out, compute_type, alg));
}

static ffdas_error_t das_sparse_call(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                               float wavenum, ffdas_tensor_desc_t x_desc, const void *x,
                               const float *dstpos, const float *offsets, const float *weights,
                               int sparse_count, const int *sparse_indices, const void *beta,
                               ffdas_tensor_desc_t out_desc, void *out,
                               ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    return das_forward_common(handle, srcpos, srcdir, wavenum, 1.0f, x_desc, x, dstpos, offsets,
                              weights, true, sparse_count, sparse_indices, beta, out_desc, out,
                              compute_type, alg);
}

ffdas_error_t ffdas_das_sparse(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                               float wavenum, ffdas_tensor_desc_t x_desc, const void *x,
                               const float *dstpos, const float *offsets, const float *weights,
                               int sparse_count, const int *sparse_indices, const void *beta,
                               ffdas_tensor_desc_t out_desc, void *out,
                               ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    ffdas::CallScope cs(handle, "ffdas_das_sparse");
    return cs.done(das_sparse_call(handle, srcpos, srcdir, wavenum, x_desc, x, dstpos, offsets,
                                   weights, sparse_count, sparse_indices, beta, out_desc, out,
                                   compute_type, alg));
}

static ffdas_error_t das_ex_call(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                           float wavenum, float scale, ffdas_tensor_desc_t x_desc, const void *x,
                           const float *dstpos, const float *offsets, const float *weights,
                           int sparse_count, const int *sparse_indices, const void *beta,
                           ffdas_tensor_desc_t out_desc, void *out,
                           ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    if (!sparse_indices && sparse_count != 0) return FFDAS_ERROR_INVALID_ARGUMENT;
    return das_forward_common(handle, srcpos, srcdir, wavenum, scale, x_desc, x, dstpos, offsets,
                              weights, sparse_indices != nullptr, sparse_count, sparse_indices,
                              beta, out_desc, out, compute_type, alg);
}

ffdas_error_t ffdas_das_ex(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                           float wavenum, float scale, ffdas_tensor_desc_t x_desc, const void *x,
                           const float *dstpos, const float *offsets, const float *weights,
                           int sparse_count, const int *sparse_indices, const void *beta,
                           ffdas_tensor_desc_t out_desc, void *out,
                           ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    ffdas::CallScope cs(handle, "ffdas_das_ex");
    return cs.done(das_ex_call(handle, srcpos, srcdir, wavenum, scale, x_desc, x, dstpos, offsets,
                               weights, sparse_count, sparse_indices, beta, out_desc, out,
                               compute_type, alg));
}

static ffdas_error_t das_backward_call(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                                 float wavenum, float scale, ffdas_tensor_desc_t x_desc,
                                 const void *x, const float *dstpos, const float *offsets,
                                 const float *weights, int sparse_count,
                                 const int *sparse_indices, ffdas_tensor_desc_t grad_out_desc,
                                 const void *grad_out, ffdas_tensor_desc_t grad_x_desc,
                                 void *grad_x, float *grad_offsets, float *grad_weights,
                                 float *grad_scale, ffdas_compute_type_t compute_type,
                                 ffdas_alg_t alg) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    FFDAS_TRY(handle->check_device());
    if (!x_desc || !grad_out_desc) return FFDAS_ERROR_NULL_POINTER;
    if (!sparse_indices && sparse_count != 0) return FFDAS_ERROR_INVALID_ARGUMENT;
    const ffdas_tensor_desc xd = *x_desc, gd = *grad_out_desc;
    DasProblem p;
    FFDAS_TRY(plan_das(xd, gd, srcpos, dstpos, offsets, weights, sparse_indices != nullptr,
                       sparse_count, sparse_indices, compute_type, alg, p));
    resolve_tune(*handle, p, srcdir != nullptr);
    p.fi_frames = fi_frames_for(p, alg);
    if (!x || !grad_out) return FFDAS_ERROR_NULL_POINTER;
    if (!(scale > 0.0f) || !std::isfinite(scale)) return FFDAS_ERROR_INVALID_ARGUMENT;
    if (grad_x) {
        if (!grad_x_desc) return FFDAS_ERROR_NULL_POINTER;
        const ffdas_tensor_desc &xg = *grad_x_desc;
        const ffdas_datatype_t want = p.acc64 ? (p.cplx ? FFDAS_C_64F : FFDAS_R_64F)
                                              : (p.cplx ? FFDAS_C_32F : FFDAS_R_32F);
        if (xg.dtype != want) return FFDAS_ERROR_UNSUPPORTED_TYPE;
        if (xg.dims != xd.dims) return FFDAS_ERROR_INVALID_DIMS;
        int64_t s = 1;
        for (int i = (int)xg.dims.size() - 1; i >= 0; i--) {
            if (xg.dims[i] > 1 && xg.strides[i] != s) return FFDAS_ERROR_NON_CONTIGUOUS;
            s *= xg.dims[i];
        }
    }
    // grad_x alone does not read x: no converted copy is needed then
    const bool need_xv = grad_offsets || grad_weights || grad_scale;
    const size_t xb = need_xv ? cast_bytes(p) : 0;
    const size_t gb = grad_x ? align_up(backward_workspace(p)) : 0;
    if (grad_x && gb == 0 && !p.plain) return FFDAS_ERROR_DIMS_TOO_LARGE;   // row exceeds LDS
    const size_t pb = align_up(param_workspace(p, grad_offsets != nullptr,
                                               grad_weights != nullptr, grad_scale != nullptr));
    void *ws = nullptr;
    FFDAS_TRY(handle->workspace(xb + gb + pb, &ws));
    XView xv;
    if (xb == 0) handle->xcopy.valid = false;   // grad_x's buffers start at the workspace base
    if (need_xv) {
        if (handle->reuse_x && xb > 0 && xcopy_matches(*handle, xd, p, x, ws)) {
            xv = copy_view(p, ws);
        } else {
            FFDAS_TRY(prepare_x(*handle, p, x, ws, xv));
            record_xcopy(*handle, xd, p, x, ws);
        }
    }
    void *gt = grad_x ? static_cast<char *>(ws) + xb : nullptr;
    void *pws = static_cast<char *>(ws) + xb + gb;
    DasGeomArgs a;
    a.srcpos = srcpos; a.srcdir = srcdir; a.dstpos = dstpos; a.offsets = offsets;
    a.weights = weights; a.sparse_indices = sparse_indices; a.wavenum = wavenum; a.scale = scale;
    return launch_backward(*handle, p, a, xv, grad_out, p.os_b, grad_x, grad_offsets,
                           grad_weights, grad_scale, gt, pws);
}

ffdas_error_t ffdas_das_backward(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                                 float wavenum, float scale, ffdas_tensor_desc_t x_desc,
                                 const void *x, const float *dstpos, const float *offsets,
                                 const float *weights, int sparse_count,
                                 const int *sparse_indices, ffdas_tensor_desc_t grad_out_desc,
                                 const void *grad_out, ffdas_tensor_desc_t grad_x_desc,
                                 void *grad_x, float *grad_offsets, float *grad_weights,
                                 float *grad_scale, ffdas_compute_type_t compute_type,
                                 ffdas_alg_t alg) {
    ffdas::CallScope cs(handle, "ffdas_das_backward");
    return cs.done(das_backward_call(handle, srcpos, srcdir, wavenum, scale, x_desc, x, dstpos,
                                     offsets, weights, sparse_count, sparse_indices,
                                     grad_out_desc, grad_out, grad_x_desc, grad_x, grad_offsets,
                                     grad_weights, grad_scale, compute_type, alg));
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// batched forward
// ---------------------------------------------------------------------------------------
static ffdas_error_t das_batched_call(ffdas_handle_t handle, int count,
                                          const ffdas_das_problem_t *problems,
                                          ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    FFDAS_TRY(handle->check_device());
    if (count < 0) return FFDAS_ERROR_INVALID_ARGUMENT;
    if (count == 0) return FFDAS_SUCCESS;
    if (!problems) return FFDAS_ERROR_NULL_POINTER;
    std::vector<BatchMember> ms(count);
    std::vector<size_t> wsoff(count);
    size_t total = 0;
    for (int i = 0; i < count; i++) {
        const ffdas_das_problem_t &q = problems[i];
        if (!q.x_desc || !q.out_desc) return FFDAS_ERROR_NULL_POINTER;
        if (!q.sparse_indices && q.sparse_count != 0) return FFDAS_ERROR_INVALID_ARGUMENT;
        BatchMember &m = ms[i];
        FFDAS_TRY(plan_das(*q.x_desc, *q.out_desc, q.srcpos, q.dstpos, q.offsets, q.weights,
                           q.sparse_indices != nullptr, q.sparse_count, q.sparse_indices,
                           compute_type, alg, m.p));
        resolve_tune(*handle, m.p, q.srcdir != nullptr);
        m.p.fi_frames = fi_frames_for(m.p, alg);
        if (!q.x || !q.out) return FFDAS_ERROR_NULL_POINTER;
        if (!(q.scale > 0.0f) || !std::isfinite(q.scale)) return FFDAS_ERROR_INVALID_ARGUMENT;
        read_beta(q.beta, m.p, m.beta_re, m.beta_im);
        m.a.srcpos = q.srcpos; m.a.srcdir = q.srcdir; m.a.dstpos = q.dstpos;
        m.a.offsets = q.offsets; m.a.weights = q.weights; m.a.sparse_indices = q.sparse_indices;
        m.a.wavenum = q.wavenum; m.a.scale = q.scale;
        m.out = q.out;
        wsoff[i] = total;
        total += cast_bytes(m.p);
    }
    void *ws = nullptr;
    FFDAS_TRY(handle->workspace(total, &ws));
    handle->xcopy.valid = false;
    // converted copies of x, then one gather launch per (storage, complex, chunk) group
    std::vector<XView> xv(count);
    for (int i = 0; i < count; i++) {
        FFDAS_TRY(prepare_x(*handle, ms[i].p, problems[i].x,
                            static_cast<char *>(ws) + wsoff[i], xv[i]));
        ms[i].xfi = xv[i].ptr;
    }
    std::vector<bool> done(count, false);
    for (int i = 0; i < count; i++) {
        if (done[i]) continue;
        if (ms[i].p.fi_frames <= 0) {
            FFDAS_TRY(launch_forward(*handle, ms[i].p, ms[i].a, xv[i], ms[i].out, ms[i].beta_re,
                                     ms[i].beta_im));
            done[i] = true;
            continue;
        }
        std::vector<BatchMember> group;
        for (int j = i; j < count; j++) {
            // one launch per storage type, complexness, frame chunk and layout (sparse
            // problems read whole frame rows, dense ones frame-group planes)
            if (done[j] || ms[j].p.fi_frames != ms[i].p.fi_frames ||
                ms[j].p.k_storage != ms[i].p.k_storage || ms[j].p.cplx != ms[i].p.cplx ||
                ms[j].p.sparse != ms[i].p.sparse)
                continue;
            group.push_back(ms[j]);
            done[j] = true;
        }
        FFDAS_TRY(launch_forward_batched(*handle, group));
    }
    return FFDAS_SUCCESS;
}

extern "C" ffdas_error_t ffdas_das_batched(ffdas_handle_t handle, int count,
                                          const ffdas_das_problem_t *problems,
                                          ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    ffdas::CallScope cs(handle, "ffdas_das_batched");
    return cs.done(das_batched_call(handle, count, problems, compute_type, alg));
}

// ---------------------------------------------------------------------------------------
// autotuner
// ---------------------------------------------------------------------------------------
namespace ffdas {
namespace {

// Average duration [ms] of `reps` runs of fn on the handle's stream, after one warm-up run.
template <class Fn>
ffdas_error_t time_runs(ffdas_context &h, int reps, Fn fn, float &ms) {
    FFDAS_TRY(fn());
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return FFDAS_ERROR_CUDA_RUNTIME;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return FFDAS_ERROR_CUDA_RUNTIME;
    }
    ffdas_error_t err = FFDAS_SUCCESS;
    hipError_t he = hipEventRecord(a, h.stream);
    for (int i = 0; i < reps && err == FFDAS_SUCCESS && he == hipSuccess; i++) err = fn();
    if (he == hipSuccess) he = hipEventRecord(b, h.stream);
    if (he == hipSuccess) he = hipEventSynchronize(b);
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (err != FFDAS_SUCCESS) return err;
    if (he != hipSuccess) return hip_status(he);
    ms /= (float)reps;
    return FFDAS_SUCCESS;
}

// Restores the handle's tuning override on scope exit.
struct ForceGuard {
    ffdas_context *h;
    explicit ForceGuard(ffdas_context *hh) : h(hh) {}
    ~ForceGuard() {
        h->force_tune = nullptr;
        h->xcopy.valid = false;
    }
};

}  // namespace
}  // namespace ffdas

static ffdas_error_t das_autotune_call(
    ffdas_handle_t handle, const float *srcpos, const float *srcdir, float wavenum, float scale,
    ffdas_tensor_desc_t x_desc, const void *x, const float *dstpos, const float *offsets,
    const float *weights, int sparse_count, const int *sparse_indices,
    ffdas_tensor_desc_t out_desc, void *out, ffdas_tensor_desc_t grad_x_desc, void *grad_x,
    ffdas_compute_type_t compute_type, ffdas_alg_t alg, int reps, ffdas_das_tuning_t *result) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    FFDAS_TRY(handle->check_device());
    if (!x_desc || !out_desc) return FFDAS_ERROR_NULL_POINTER;
    if (!x || !out) return FFDAS_ERROR_NULL_POINTER;
    if (!sparse_indices && sparse_count != 0) return FFDAS_ERROR_INVALID_ARGUMENT;
    reps = reps < 1 ? 1 : (reps > 100 ? 100 : reps);
    const bool sparse = sparse_indices != nullptr;
    DasProblem p;
    FFDAS_TRY(plan_das(*x_desc, *out_desc, srcpos, dstpos, offsets, weights, sparse,
                       sparse_count, sparse_indices, compute_type, alg, p));
    const std::string key = tune_key(*handle, p, srcdir != nullptr);
    ForceGuard guard(handle);
    handle->xcopy.valid = false;
    const double zero[2] = {0.0, 0.0};   // host beta of any output type
    auto forward = [&]() {
        return das_forward_common(handle, srcpos, srcdir, wavenum, scale, x_desc, x, dstpos,
                                  offsets, weights, sparse, sparse_count, sparse_indices, zero,
                                  out_desc, out, compute_type, alg);
    };
    int candidates = 0;
    auto measure = [&](const DasTune &t, auto fn, float &ms) {
        handle->force_tune = &t;
        candidates++;
        return time_runs(*handle, reps, fn, ms);
    };

    // The frame chunking and the 2-D tile shape are shared by the forward and the parameter
    // adjoint (one layout of x), so with the adjoint they are chosen on the sum of both
    // times; the channel groups of each kernel are chosen separately for every such pair.
    const DasTune heuristic;
    float fwd_default = 0.0f;
    FFDAS_TRY(measure(heuristic, forward, fwd_default));
    std::vector<int> frames;
    for (int F : {16, 8, 4}) {
        DasProblem q = p;
        q.tune.fi_frames = F;
        if (fi_frames_for(q, alg) == F) frames.push_back(F);
    }
    int dims2 = 0;
    for (int64_t v : p.spatial) dims2 += v > 1 ? 1 : 0;
    std::vector<std::pair<int, int>> tiles = {{0, 0}};
    if (dims2 == 2) tiles = {{8, 8}, {16, 4}, {4, 16}, {32, 2}};
    std::vector<int> groups = {0};
    if (!frames.empty()) {
        groups.clear();
        for (int G : {1, 2, 4, 8})
            if (G <= p.C) groups.push_back(G);
    } else {
        frames.push_back(0);
    }

    float *gofs = nullptr;
    if (grad_x) {
        const size_t gbytes = (size_t)p.K * (size_t)p.N * sizeof(float);
        if (hipMalloc(reinterpret_cast<void **>(&gofs), gbytes) != hipSuccess)
            return FFDAS_ERROR_ALLOCATION_FAILED;
    }
    struct FreeOnExit {
        float *p;
        ~FreeOnExit() { if (p) (void)hipFree(p); }
    } free_gofs{gofs};
    auto param = [&]() {
        return ffdas_das_backward(handle, srcpos, srcdir, wavenum, scale, x_desc, x, dstpos,
                                  offsets, weights, sparse_count, sparse_indices, out_desc, out,
                                  nullptr, nullptr, gofs, nullptr, nullptr, compute_type, alg);
    };
    auto bwdx = [&]() {
        return ffdas_das_backward(handle, srcpos, srcdir, wavenum, scale, x_desc, x, dstpos,
                                  offsets, weights, sparse_count, sparse_indices, out_desc, out,
                                  grad_x_desc, grad_x, nullptr, nullptr, nullptr, compute_type,
                                  alg);
    };

    DasTune best = heuristic;
    float fwd_best = fwd_default, param_best = 0.0f, param_default = 0.0f;
    float total_best = 0.0f;
    bool have = false;
    if (grad_x) {
        FFDAS_TRY(forward());   // out holds a forward result: the grad_out of the timed adjoints
        FFDAS_TRY(measure(heuristic, param, param_default));
    }
    for (int F : frames)
        for (auto tt : tiles) {
            DasTune layout = heuristic;
            layout.fi_frames = F;
            layout.tile_h = tt.first;
            layout.tile_w = tt.second;
            DasTune bf = layout;
            float fb = 0.0f;
            for (int G : groups) {
                DasTune t = layout;
                t.groups = G;
                float ms = 0.0f;
                FFDAS_TRY(measure(t, forward, ms));
                if (G == groups.front() || ms < fb) {
                    fb = ms;
                    bf = t;
                }
            }
            float pb = 0.0f;
            if (grad_x) {
                for (int G : groups) {
                    DasTune t = bf;
                    t.param_groups = G;
                    float ms = 0.0f;
                    FFDAS_TRY(measure(t, param, ms));
                    if (G == groups.front() || ms < pb) {
                        pb = ms;
                        bf.param_groups = G;
                    }
                }
            }
            if (!have || fb + pb < total_best) {
                have = true;
                total_best = fb + pb;
                best = bf;
                fwd_best = fb;
                param_best = pb;
            }
        }
    if (fwd_default + param_default <= total_best) {   // the heuristic is at least as fast
        best = heuristic;
        fwd_best = fwd_default;
        param_best = param_default;
    }

    // grad_x: frames per workgroup, lanes along depth, workgroup size
    float bwd_default = 0.0f, bwd_best = 0.0f;
    if (grad_x && !p.plain) {
        DasTune t0 = best;
        t0.bwdx_frames = 0;
        t0.bwdx_lg = -1;
        t0.bwdx_threads = 0;
        float xd = 0.0f;
        FFDAS_TRY(measure(t0, bwdx, xd));
        float xb = xd;
        DasTune bx = t0;
        for (int F : {16, 8, 4, 2})
            for (int lg : {2, 3, 4, 5, 6})
                for (int th : {512, 1024}) {
                    DasTune t = t0;
                    t.bwdx_frames = F;
                    t.bwdx_lg = lg;
                    t.bwdx_threads = th;
                    DasProblem q = p;
                    q.tune = t;
                    if (bwdx_frames(q) != F) continue;
                    float ms = 0.0f;
                    FFDAS_TRY(measure(t, bwdx, ms));
                    if (ms < xb) {
                        xb = ms;
                        bx = t;
                    }
                }
        best.bwdx_frames = bx.bwdx_frames;
        best.bwdx_lg = bx.bwdx_lg;
        best.bwdx_threads = bx.bwdx_threads;
        bwd_default = param_default + xd;
        bwd_best = param_best + xb;
    } else if (grad_x) {
        float xd = 0.0f;
        FFDAS_TRY(measure(best, bwdx, xd));
        bwd_default = param_default + xd;
        bwd_best = param_best + xd;
    }
    tune_store(key, best);
    if (result) {
        result->groups = best.groups;
        result->fi_frames = best.fi_frames;
        result->tile_h = best.tile_h;
        result->tile_w = best.tile_w;
        result->param_groups = best.param_groups;
        result->bwdx_frames = best.bwdx_frames;
        result->bwdx_depth_lanes = best.bwdx_lg >= 0 ? (1 << best.bwdx_lg) : 0;
        result->bwdx_threads = best.bwdx_threads;
        result->forward_ms = fwd_best;
        result->backward_ms = bwd_best;
        result->default_forward_ms = fwd_default;
        result->default_backward_ms = bwd_default;
        result->candidates = candidates;
    }
    return FFDAS_SUCCESS;
}

extern "C" ffdas_error_t ffdas_das_autotune(
    ffdas_handle_t handle, const float *srcpos, const float *srcdir, float wavenum, float scale,
    ffdas_tensor_desc_t x_desc, const void *x, const float *dstpos, const float *offsets,
    const float *weights, int sparse_count, const int *sparse_indices,
    ffdas_tensor_desc_t out_desc, void *out, ffdas_tensor_desc_t grad_x_desc, void *grad_x,
    ffdas_compute_type_t compute_type, ffdas_alg_t alg, int reps, ffdas_das_tuning_t *result) {
    ffdas::CallScope cs(handle, "ffdas_das_autotune");
    return cs.done(das_autotune_call(handle, srcpos, srcdir, wavenum, scale, x_desc, x, dstpos,
                                     offsets, weights, sparse_count, sparse_indices, out_desc,
                                     out, grad_x_desc, grad_x, compute_type, alg, reps, result));
}

extern "C" ffdas_error_t ffdas_set_reuse_x(ffdas_handle_t handle, int enable) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    handle->reuse_x = enable != 0;
    return FFDAS_SUCCESS;
}
