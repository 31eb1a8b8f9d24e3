// Internal types of libffdas.so (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "ffdas.h"

// Tensor descriptor: dims/strides in elements (complex = one element).
struct ffdas_tensor_desc {
    std::vector<int64_t> dims;
    std::vector<int64_t> strides;
    ffdas_datatype_t dtype = FFDAS_R_32F;

    // product of the dims, saturated at INT64_MAX (such a tensor is rejected elsewhere)
    int64_t numel() const {
        int64_t n = 1;
        for (int64_t d : dims)
            if (__builtin_mul_overflow(n, d, &n)) return INT64_MAX;
        return n;
    }
};

// Library handle: device, stream, workspace.
namespace ffdas {
// Launch parameters a tuner may choose; 0 (-1 for bwdx_lg) selects the built-in heuristic.
struct DasTune {
    int groups = 0;         // waves per voxel tile splitting the channels (forward gather)
    int fi_frames = 0;      // frames per frame-inner chunk
    int tile_h = 0;         // 2-D wave tile, tile_h * tile_w == 64
    int tile_w = 0;
    int param_groups = 0;   // waves per voxel tile in the parameter adjoint
    int bwdx_frames = 0;    // frames per grad_x workgroup
    int bwdx_lg = -1;       // log2 of the lanes along depth in the grad_x scatter
    int bwdx_threads = 0;   // grad_x workgroup size
};
}  // namespace ffdas

struct ffdas_context {
    int device_id = 0;
    int cu_count = 0;
    int lds_bytes = 0;
    std::string arch;          // e.g. "gfx950"
    hipStream_t stream = nullptr;
    // stream-ordered workspace (grow-only, released with hipFreeAsync on the owning stream)
    void *ws = nullptr;
    size_t ws_bytes = 0;
    hipStream_t ws_stream = nullptr;
    // earlier, smaller workspace blocks: kept until the handle is destroyed, because a graph
    // captured while one of them was current still reads it on every replay
    std::vector<void *> retired;

    // set by the autotuner while it times candidates: replaces the cached/heuristic choice
    const ffdas::DasTune *force_tune = nullptr;
    // The converted copy of x at the start of the workspace: which x it was made from.
    // ffdas_das_backward reuses it when reuse_x is set and everything matches.
    struct {
        bool valid = false;
        const void *x = nullptr;
        void *ws = nullptr;
        std::vector<int64_t> dims, strides;
        int dtype = 0, storage = 0, fi_frames = 0, cast = 0, sparse = 0;
    } xcopy;
    bool reuse_x = false;

    // rocBLAS handle used by the rank-truncation eigensolver (created on first use)
    void *blas = nullptr;
    // pinned host word into which truncate_rank reads its solver status (allocated on first
    // use: a copy into pageable memory goes through a staging buffer)
    int *status_host = nullptr;

    // A second stream (default priority, non-blocking) on which ffdas_das_backward runs the
    // grad_x scatter beside the parameter adjoint, and the events that fork it from and join
    // it back into `stream`; created on first use.
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    ffdas_error_t side_stream(hipStream_t *s);

    ffdas_error_t check_device() const;
    ffdas_error_t workspace(size_t bytes, void **ptr);
};

namespace ffdas {

inline bool is_complex(ffdas_datatype_t t) {
    return t == FFDAS_C_16F || t == FFDAS_C_32F || t == FFDAS_C_64F || t == FFDAS_C_16BF ||
           t == FFDAS_C_32I || t == FFDAS_C_16I;
}

// Storage classes understood by the kernels.
enum class Storage : int { F16 = 0, BF16 = 1, F32 = 2, F64 = 3 };

// Decoded problem (see das_plan.cpp).
struct DasProblem {
    int B = 1, C = 0, T = 0, S = 0, K = 0;
    int64_t N = 0;
    bool have_batch = false;
    bool cplx = false;
    bool sparse = false;
    Storage x_storage = Storage::F32;     // storage of the input as given
    Storage k_storage = Storage::F32;     // storage the kernel gathers from
    bool acc64 = false;                   // double accumulation (double data)
    bool x_needs_copy = false;            // cast / relayout into a contiguous workspace
    int fi_frames = 0;                    // > 0: frame-inner chunks of this many frames
    bool plain = false;                   // ALG1: plain kernels (natural layout, atomic adjoint)
    DasTune tune;                         // launch parameters (tuned or heuristic)
    int64_t xs_b = 0, xs_c = 0, xs_t = 0; // strides of x as given (elements)
    int64_t os_b = 0;                     // batch stride of out (elements)
    // spatial shape of the output (without batch), used for voxel tiling
    std::vector<int64_t> spatial;
};

ffdas_error_t plan_das(const ffdas_tensor_desc &x, const ffdas_tensor_desc &out,
                       const float *srcpos, const float *dstpos, const float *offsets,
                       const float *weights, bool sparse_mode, int sparse_count,
                       const int *sparse_indices, ffdas_compute_type_t ct, ffdas_alg_t alg,
                       DasProblem &p);

ffdas_error_t hip_status(hipError_t e, ffdas_error_t kind = FFDAS_ERROR_CUDA_RUNTIME);

}  // namespace ffdas

namespace ffdas {
// Scope of one C-ABI call: a named roctx range (if the roctx library is present and
// FFDAS_ROCTX is not 0) and, with FFDAS_DEBUG set, done() waits for the handle's stream and
// turns an asynchronous failure into FFDAS_ERROR_CUDA_RUNTIME.
class CallScope {
  public:
    CallScope(const ffdas_context *h, const char *name);
    ~CallScope();
    CallScope(const CallScope &) = delete;
    CallScope &operator=(const CallScope &) = delete;
    ffdas_error_t done(ffdas_error_t status) const;

  private:
    const ffdas_context *h_;
    bool pushed_ = false;
};
bool debug_sync_enabled();
}  // namespace ffdas

#define FFDAS_TRY(expr)                               \
    do {                                              \
        ffdas_error_t _e = (expr);                    \
        if (_e != FFDAS_SUCCESS) return _e;           \
    } while (0)

namespace ffdas {

// Geometry pointers and scalars of one call (device pointers).
struct DasGeomArgs {
    const float *srcpos = nullptr;
    const float *srcdir = nullptr;
    const float *dstpos = nullptr;
    const float *offsets = nullptr;
    const float *weights = nullptr;
    const int *sparse_indices = nullptr;
    float wavenum = 0.0f;
    float scale = 1.0f;
};

// Channel data as the kernels see it: storage p.k_storage, strides in elements.
struct XView {
    const void *ptr = nullptr;
    int64_t sb = 0, sc = 0, st = 0;
};

ffdas_error_t launch_forward(ffdas_context &h, const DasProblem &p, const DasGeomArgs &a,
                             const XView &xv, void *out, double beta_re, double beta_im);
ffdas_error_t launch_backward(ffdas_context &h, const DasProblem &p, const DasGeomArgs &a,
                              const XView &xv, const void *gout, int64_t gs_b, void *gx,
                              float *gofs, float *gwt, float *gscale, void *gt, void *pws);
// workspace of the parameter adjoint's partial sums (frame chunks, d/d scale per block)
size_t param_workspace(const DasProblem &p, bool has_offsets, bool has_weights, bool has_scale);
// bytes of workspace the grad_x scatter needs for the transposed grad_out
size_t backward_workspace(const DasProblem &p);
int bwdx_frames(const DasProblem &p);
ffdas_error_t launch_cast(ffdas_context &h, const DasProblem &p, const void *x, void *dst);
ffdas_error_t launch_fill_zero(ffdas_context &h, float *p, int64_t n);

size_t storage_bytes(Storage s);

// opt-in per-kernel timing (profile.cpp)
bool profiling_enabled();
void *profile_begin(hipStream_t st);
void profile_end(void *begin, hipStream_t st, const char *name);
struct ProfScope {
    void *ev;
    hipStream_t st;
    const char *name;
    ProfScope(hipStream_t s, const char *n) : ev(profile_begin(s)), st(s), name(n) {}
    ~ProfScope() { profile_end(ev, st, name); }
};
int fi_frames_for(const DasProblem &p, ffdas_alg_t alg);

// One problem of a batched forward launch: planned problem, geometry, its frame-inner copy
// of x, output and beta.
struct BatchMember {
    DasProblem p;
    DasGeomArgs a;
    const void *xfi = nullptr;
    void *out = nullptr;
    double beta_re = 0.0, beta_im = 0.0;
};
// All members share storage type, complexness and frames per chunk (p.fi_frames > 0).
ffdas_error_t launch_forward_batched(ffdas_context &h, const std::vector<BatchMember> &ms);

// The forward of ffdas_das / ffdas_das_sparse / ffdas_das_ex (das_api.cpp) with its workspace
// (the converted copy of x) taken from get_ws(bytes, &ptr); record_copy marks that copy for
// reuse by a later ffdas_das_backward on the same handle.
using WorkspaceFn = std::function<ffdas_error_t(size_t, void **)>;
ffdas_error_t das_forward_with(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                               float wavenum, float scale, ffdas_tensor_desc_t x_desc,
                               const void *x, const float *dstpos, const float *offsets,
                               const float *weights, bool sparse_mode, int sparse_count,
                               const int *sparse_indices, const void *beta,
                               ffdas_tensor_desc_t out_desc, void *out,
                               ffdas_compute_type_t compute_type, ffdas_alg_t alg,
                               const WorkspaceFn &get_ws, bool record_copy);

// tuning cache (tune.cpp)
std::string tune_key(const ffdas_context &h, const DasProblem &p, bool has_srcdir);
void resolve_tune(const ffdas_context &h, DasProblem &p, bool has_srcdir);
void tune_store(const std::string &key, const DasTune &t);
size_t fi_bytes(const DasProblem &p);
ffdas_error_t launch_relayout_fi(ffdas_context &h, const DasProblem &p, const void *x, void *dst);

}  // namespace ffdas

namespace ffdas {

// One ffdas_truncate_rank call, as the kernels see it (rank_kernels.hip).  Element strides
// are in elements (complex = one element).
struct RankPlan {
    bool dbl = false, cplx = false;
    int64_t m = 0, n = 0;          // x is (m, n)
    int p = 0;                     // min(m, n): Gram size
    int64_t L = 0;                 // max(m, n): Gram vector length = projection vector count
    int64_t gsa = 0, gsk = 0;      // Gram vector a, element k at a*gsa + k*gsk
    bool gvec = false;             // 16-byte loads along k
    int64_t xsj = 0, xsf = 0;      // projection vector j, element f of x at j*xsj + f*xsf
    int64_t osj = 0, osf = 0;      // same for out
    bool xvec = false, ovec = false;
    bool keep = true;              // project onto the kept set (else subtract the dropped set)
    bool lds_gram = false;         // LDS-staged Gram kernel (FP32, p <= 256)
    int r = 0;                     // eigenvector columns used
    int q_split = 0, c_first = 0, c_second = 0;
    // launch shape (rank_plan_launch_shape)
    int bs = 0, nb = 0, splits = 0, pp = 0, rp = 0, rt = 0, passes = 0;
    int64_t kchunk = 0;
    size_t part_bytes = 0;
};

void rank_plan_launch_shape(RankPlan &pl, int cu_count);
ffdas_error_t launch_rank_gram(hipStream_t st, const RankPlan &pl, const void *x, void *part,
                               void *H);
bool rank_jacobi_fits(const RankPlan &pl, int lds_bytes);
ffdas_error_t launch_rank_jacobi(hipStream_t st, const RankPlan &pl, void *H, int *info,
                                 int max_sweeps);
// FP32, 2 <= p <= 224 real / 160 complex: tridiagonalisation (D, E of the real tridiagonal,
// packed reflectors U and tau; clears stat[0..7]), the tridiagonal's eigenvectors for the
// plan's r columns (Y, real p x r), and their back-transformation into V's columns.
bool rank_tridiag_fits(const RankPlan &pl, int lds_bytes);
ffdas_error_t launch_rank_tridiag(hipStream_t st, const RankPlan &pl, int lds_bytes,
                                  const void *H, void *D, void *E, void *U, void *tau,
                                  int *stat);
ffdas_error_t launch_rank_tridiag_eigvec(hipStream_t st, const RankPlan &pl, const void *D,
                                         const void *E, void *Y, int *stat);
ffdas_error_t launch_rank_backtransform(hipStream_t st, const RankPlan &pl, int lds_bytes,
                                        const void *U, const void *tau, const void *Y,
                                        const int *stat, void *V, int *info);
ffdas_error_t launch_rank_project(hipStream_t st, const RankPlan &pl, const void *V,
                                  const void *x, void *out, void *Qh, void *Qn);
ffdas_error_t launch_rank_copy(hipStream_t st, bool dbl, bool cplx, const void *x, int64_t xs0,
                               int64_t xs1, void *out, int64_t os0, int64_t os1, int64_t m,
                               int64_t n);
void release_rank_solver(ffdas_context &h);

}  // namespace ffdas
