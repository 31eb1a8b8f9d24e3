/*
 * ffdas C-ABI for the MI355X (gfx950) delay-and-sum library, libffdas.so.
 *
 * Drop-in boundary: every enum value and every function signature that the reference
 * declares for the delay-and-sum path is kept bit-for-bit, so callers written against the
 * reference header (Python ctypes/nanobind, MATLAB MEX, C/C++) link unchanged.  Each
 * declaration cites the reference declaration it replaces (include/ffdas.h:<line> in the
 * reference tree).  Pointers named "device" must point to memory on the handle's GPU;
 * everything runs asynchronously on the handle's stream and no call synchronises.
 *
 * From the reference's other operations, truncate_rank, interpolation and greens are declared
 * here with their reference signatures; contraction, gather and scatter are not part of this
 * library.
 *
 * Additive extensions (new symbols or new enum values; no existing value changes meaning)
 * are marked "EXTENSION".
 */
#ifndef FFDAS_H_
#define FFDAS_H_

#define FFDAS_API __attribute__((visibility("default")))

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

/* Error codes; values as in the reference (include/ffdas.h:19-62).  The CUDA-named codes
 * report the corresponding HIP runtime failures. */
typedef enum {
    FFDAS_SUCCESS = 0,
    FFDAS_ERROR_INVALID_ARGUMENT = 1,
    FFDAS_ERROR_INVALID_DIMS = 2,
    FFDAS_ERROR_UNSUPPORTED_TYPE = 3,
    FFDAS_ERROR_NULL_POINTER = 4,
    FFDAS_ERROR_INVALID_HANDLE = 5,
    FFDAS_ERROR_NON_CONTIGUOUS = 6,
    FFDAS_ERROR_DIMS_TOO_LARGE = 7,
    FFDAS_ERROR_ALLOCATION_FAILED = 10,
    FFDAS_ERROR_DEVICE_MEMORY_INSUFFICIENT = 11,
    FFDAS_ERROR_HOST_MEMORY_INSUFFICIENT = 12,
    FFDAS_ERROR_INVALID_DEVICE = 13,
    FFDAS_ERROR_DEVICE_NOT_SUPPORTED = 20,
    FFDAS_ERROR_INSUFFICIENT_COMPUTE_CAPABILITY = 21,
    FFDAS_ERROR_DEVICE_NOT_AVAILABLE = 22,
    FFDAS_ERROR_DEVICE_MISMATCH = 23,
    FFDAS_ERROR_CUDA_DRIVER = 30,
    FFDAS_ERROR_CUDA_RUNTIME = 31,
    FFDAS_ERROR_CUDA_LAUNCH_FAILED = 32,
    FFDAS_ERROR_CUDA_SYNC_FAILED = 33,
    FFDAS_ERROR_CUDA_MEMCPY_FAILED = 34,
    FFDAS_ERROR_CUBLAS = 40,
    FFDAS_ERROR_CUSOLVER = 41,
    FFDAS_ERROR_CUFFT = 42,
    FFDAS_ERROR_NVTX = 43,
    FFDAS_ERROR_ALGORITHM_FAILED = 50,
    FFDAS_ERROR_CONVERGENCE_FAILED = 51,
    FFDAS_ERROR_NUMERICAL_INSTABILITY = 52,
    FFDAS_ERROR_NOT_INITIALIZED = 60,
    FFDAS_ERROR_ALREADY_INITIALIZED = 61,
    FFDAS_ERROR_INVALID_STATE = 62,
    FFDAS_ERROR_OUT_OF_BOUNDS_INDEX = 63,
    FFDAS_ERROR_INTERNAL = 100,
    FFDAS_ERROR_UNKNOWN = 101
} ffdas_error_t;

/* Element types; values as in the reference (include/ffdas.h:64-75).  Complex types are
 * interleaved (re, im) pairs and count as one element in dims and strides.
 * EXTENSION: FFDAS_R_16BF / FFDAS_C_16BF (bfloat16 storage) at previously unused values. */
typedef enum {
    FFDAS_R_16F = 2,
    FFDAS_C_16F = 6,
    FFDAS_R_32F = 0,
    FFDAS_C_32F = 4,
    FFDAS_R_64F = 16,
    FFDAS_C_64F = 17,
    FFDAS_R_32I = 10,
    FFDAS_C_32I = 11,
    FFDAS_R_16I = 14,
    FFDAS_C_16I = 15,
    FFDAS_R_16BF = 18,
    FFDAS_C_16BF = 19
} ffdas_datatype_t;

/* Storage precision requested for the channel data; values as in the reference
 * (include/ffdas.h:77-82).  Arithmetic is FP32 (FP64 for double data) in every case; the
 * compute type only selects how x is stored while it is gathered.
 * EXTENSION: FFDAS_COMPUTE_16BF stores float data as bfloat16. */
typedef enum {
    FFDAS_COMPUTE_DEFAULT = 0,
    FFDAS_COMPUTE_16F = 1,
    FFDAS_COMPUTE_32F = 2,
    FFDAS_COMPUTE_64F = 3,
    FFDAS_COMPUTE_16BF = 4
} ffdas_compute_type_t;

/* Algorithm selector; values as in the reference (include/ffdas.h:84-90).  DEFAULT, ALG1,
 * ALG2 and ALG4 all run this library's single gfx950 kernel family (configuration picked by
 * the autotuner); ALG3 is rejected with FFDAS_ERROR_INVALID_ARGUMENT as in the reference. */
typedef enum {
    FFDAS_ALG_DEFAULT = 0,
    FFDAS_ALG1 = 1,
    FFDAS_ALG2 = 2,
    FFDAS_ALG3 = 3,
    FFDAS_ALG4 = 4
} ffdas_alg_t;

/* Interpolation mode; values as in the reference (include/ffdas.h:92-95). */
typedef enum {
    FFDAS_INTERP_NEAREST = 0,
    FFDAS_INTERP_LINEAR = 1
} ffdas_interp_mode_t;

struct ffdas_context;
typedef struct ffdas_context *ffdas_handle_t;

struct ffdas_interpolation_plan;
typedef struct ffdas_interpolation_plan *ffdas_interpolation_plan_t;

struct ffdas_tensor_desc;
typedef struct ffdas_tensor_desc *ffdas_tensor_desc_t;

#if defined(__cplusplus)
extern "C" {
#endif

/* Bytes per element, 0 for an unknown type.  Replaces include/ffdas.h:115. */
FFDAS_API size_t ffdas_type_size(ffdas_datatype_t datatype);

/* Static description of an error code.  Replaces include/ffdas.h:117. */
FFDAS_API const char *ffdas_error_string(ffdas_error_t error);

/* Tensor descriptor: dims and strides in ELEMENTS, strides == NULL means C-contiguous.
 * The library copies descriptors at every call.  Replaces include/ffdas.h:119-129. */
FFDAS_API ffdas_error_t ffdas_create_tensor_desc(ffdas_tensor_desc_t *desc, int64_t ndim,
                                                 const int64_t *dims, const int64_t *strides,
                                                 ffdas_datatype_t dtype);
FFDAS_API ffdas_error_t ffdas_destroy_tensor_desc(ffdas_tensor_desc_t desc);

/* Handle bound to the device current at creation, with its own stream (initially the
 * null stream) and stream-ordered workspace.  Not thread-safe.  Returns
 * FFDAS_ERROR_DEVICE_NOT_AVAILABLE when no GPU is present.
 * Replaces include/ffdas.h:131-147. */
FFDAS_API ffdas_error_t ffdas_create(ffdas_handle_t *handle);
FFDAS_API ffdas_error_t ffdas_destroy(ffdas_handle_t handle);
FFDAS_API ffdas_error_t ffdas_set_stream(ffdas_handle_t handle, uintptr_t stream);
FFDAS_API ffdas_error_t ffdas_get_stream(ffdas_handle_t handle, uintptr_t *stream);

/* Strided copy with element-type conversion: x (any rank, any element strides, negative
 * included) into out (same dims, contiguous), else FFDAS_ERROR_INVALID_DIMS.  The (x, out)
 * type pairs of the reference: R16I->{R16I,R32F}, R16F->{R16F,R32F}, R32F->{R16F,R32F,R64F},
 * R64F->{R64F,R32F} and the complex counterparts C16I, C16F, C32F, C64F likewise; EXTENSION:
 * R16BF->{R16BF,R32F}, R32F->R16BF and the complex counterparts.  Other pairs return
 * FFDAS_ERROR_UNSUPPORTED_TYPE.  Narrowing rounds to nearest even.
 * Replaces include/ffdas.h:149-155. */
FFDAS_API ffdas_error_t ffdas_contiguous_copy(ffdas_handle_t handle, ffdas_tensor_desc_t x_desc,
                                              const void *x, ffdas_tensor_desc_t out_desc,
                                              void *out);

/* Delay-and-sum, compounding over all transmits.
 *   srcpos  device (C,3) f32, channel positions [sampling wavelengths]
 *   srcdir  device (C,4) f32 or NULL: unit normal + cosine of the acceptance half-angle
 *   wavenum phase rotation [rad/sample] applied to complex data, 0 disables
 *   x       device ([B,]C,T,S) per x_desc (sample stride must be 1)
 *   dstpos  device (N,3) f32 voxel positions [sampling wavelengths]
 *   offsets device (T,N) f32 transmit delays [samples]; weights device (T,N) f32
 *   beta    HOST pointer to one output-type scalar: out = beta*out + result (0: overwrite)
 *   out     device ([B,]spatial...) per out_desc, contiguous
 * Supported (x, out) pairs: (R16F,R32F) (R16BF,R32F) (R32F,R32F) (R64F,R64F)
 * (C16F,C32F) (C16BF,C32F) (C32F,C32F) (C64F,C64F).  Replaces include/ffdas.h:188-203. */
FFDAS_API ffdas_error_t ffdas_das(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                                  float wavenum, ffdas_tensor_desc_t x_desc, const void *x,
                                  const float *dstpos, const float *offsets, const float *weights,
                                  const void *beta, ffdas_tensor_desc_t out_desc, void *out,
                                  ffdas_compute_type_t compute_type, ffdas_alg_t alg);

/* Sparse compounding: every voxel compounds the sparse_count transmits listed in
 * sparse_indices (device (K,N) int32, 0-based); offsets/weights are (K,N).  An index
 * outside [0,T) contributes nothing.  Replaces include/ffdas.h:205-222. */
FFDAS_API ffdas_error_t ffdas_das_sparse(ffdas_handle_t handle, const float *srcpos,
                                         const float *srcdir, float wavenum,
                                         ffdas_tensor_desc_t x_desc, const void *x,
                                         const float *dstpos, const float *offsets,
                                         const float *weights, int sparse_count,
                                         const int *sparse_indices, const void *beta,
                                         ffdas_tensor_desc_t out_desc, void *out,
                                         ffdas_compute_type_t compute_type, ffdas_alg_t alg);

/* Device selection and events (hipEvent_t as uintptr_t; event_record uses the handle's
 * stream).  Replaces include/ffdas.h:290-324. */
FFDAS_API ffdas_error_t ffdas_device_count(int *count);
FFDAS_API ffdas_error_t ffdas_device_get(int *device);
FFDAS_API ffdas_error_t ffdas_device_set(int device);
FFDAS_API ffdas_error_t ffdas_event_create(uintptr_t *event);
FFDAS_API ffdas_error_t ffdas_event_destroy(uintptr_t event);
FFDAS_API ffdas_error_t ffdas_event_record(ffdas_handle_t handle, uintptr_t event);
FFDAS_API ffdas_error_t ffdas_event_synchronize(uintptr_t event);
FFDAS_API ffdas_error_t ffdas_event_elapsed_time(uintptr_t start, uintptr_t stop, float *ms);

/* ---------------------------------------------------------------------------------------
 * EXTENSIONS (no counterpart in the reference header)
 * ------------------------------------------------------------------------------------- */

/* Forward with a global delay scale s (tau = s * (offsets + r), s = 1 reproduces ffdas_das)
 * and optional sparse compounding (sparse_indices == NULL: dense, sparse_count must be 0).
 * Extends include/ffdas.h:188-222 for the differentiable (sound-speed) form. */
FFDAS_API ffdas_error_t ffdas_das_ex(ffdas_handle_t handle, const float *srcpos,
                                     const float *srcdir, float wavenum, float scale,
                                     ffdas_tensor_desc_t x_desc, const void *x,
                                     const float *dstpos, const float *offsets,
                                     const float *weights, int sparse_count,
                                     const int *sparse_indices, const void *beta,
                                     ffdas_tensor_desc_t out_desc, void *out,
                                     ffdas_compute_type_t compute_type, ffdas_alg_t alg);

/* Adjoint of ffdas_das_ex for grad_out (same shape/type as out).  Any output may be NULL:
 *   grad_x        device, grad_x_desc: same dims as x, contiguous, type R32F/C32F (R64F/C64F
 *                 for double data); overwritten
 *   grad_offsets  device (K,N) f32, overwritten
 *   grad_weights  device (K,N) f32, overwritten
 *   grad_scale    device, one f32, overwritten: dL/ds
 * Interpolation indices are recomputed, never stored.  Complex data follows the real-loss
 * convention grad = dL/dRe + i dL/dIm. */
FFDAS_API ffdas_error_t ffdas_das_backward(ffdas_handle_t handle, const float *srcpos,
                                           const float *srcdir, float wavenum, float scale,
                                           ffdas_tensor_desc_t x_desc, const void *x,
                                           const float *dstpos, const float *offsets,
                                           const float *weights, int sparse_count,
                                           const int *sparse_indices,
                                           ffdas_tensor_desc_t grad_out_desc,
                                           const void *grad_out, ffdas_tensor_desc_t grad_x_desc,
                                           void *grad_x, float *grad_offsets, float *grad_weights,
                                           float *grad_scale, ffdas_compute_type_t compute_type,
                                           ffdas_alg_t alg);

/* ffdas_das without device allocation: the same arguments, checks and result, with the
 * converted copy of x (FP16/BF16 cast, frame-inner relayout) kept in a device array of
 * 256 MiB that the library holds on each GPU it runs on until the process ends, instead of
 * in the handle's workspace.  No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  A problem whose copy exceeds 256 MiB returns FFDAS_ERROR_DIMS_TOO_LARGE.
 * Calls that overlap on one GPU share the array and race. */
FFDAS_API ffdas_error_t ffdas_das_noalloc(ffdas_handle_t handle, const float *srcpos,
                                          const float *srcdir, float wavenum,
                                          ffdas_tensor_desc_t x_desc, const void *x,
                                          const float *dstpos, const float *offsets,
                                          const float *weights, const void *beta,
                                          ffdas_tensor_desc_t out_desc, void *out,
                                          ffdas_compute_type_t compute_type, ffdas_alg_t alg);

/* Host-only argument check: returns the error code ffdas_das_ex would return for these
 * descriptors and scalars, without touching a device (pointers are only NULL-checked). */
FFDAS_API ffdas_error_t ffdas_das_validate(ffdas_tensor_desc_t x_desc,
                                           ffdas_tensor_desc_t out_desc, const float *srcpos,
                                           const float *dstpos, const float *offsets,
                                           const float *weights, int sparse_count,
                                           const int *sparse_indices,
                                           ffdas_compute_type_t compute_type, ffdas_alg_t alg);

/* One problem of ffdas_das_batched: the arguments of ffdas_das_ex. */
typedef struct ffdas_das_problem {
    const float *srcpos;
    const float *srcdir;
    float wavenum;
    float scale;
    ffdas_tensor_desc_t x_desc;
    const void *x;
    const float *dstpos;
    const float *offsets;
    const float *weights;
    int sparse_count;
    const int *sparse_indices;
    const void *beta;            /* host scalar of the output type */
    ffdas_tensor_desc_t out_desc;
    void *out;
} ffdas_das_problem_t;

/* Several independent forward problems (mixed probes, grids, frame ensembles, sparse
 * tables) in one call: each problem's result is that of ffdas_das_ex with the same
 * arguments.  Problems that share the kernel storage type, complexness and frame chunking
 * run in one gather launch (up to 12 problems per launch, descriptors passed by value, so a
 * graph captured around this call holds them); the others run one launch each.  Form (2)
 * of the batched/graph API; no reference counterpart (the reference launches one problem per
 * call, src/das/das.cu:9-83). */
FFDAS_API ffdas_error_t ffdas_das_batched(ffdas_handle_t handle, int count,
                                          const ffdas_das_problem_t *problems,
                                          ffdas_compute_type_t compute_type, ffdas_alg_t alg);

/* Launch parameters chosen by ffdas_das_autotune, with the measured times [ms] of the
 * chosen and of the built-in (heuristic) parameters.  0 means "heuristic" for a field. */
typedef struct ffdas_das_tuning {
    int groups;            /* waves per voxel tile splitting the channels (forward) */
    int fi_frames;         /* frames per frame-inner chunk of x */
    int tile_h, tile_w;    /* 2-D wave tile of voxels (tile_h * tile_w = 64) */
    int param_groups;      /* waves per voxel tile in the offsets/weights/scale adjoint */
    int bwdx_frames;       /* frames per grad_x workgroup */
    int bwdx_depth_lanes;  /* lanes of a wave along the first grid dimension in grad_x */
    int bwdx_threads;      /* grad_x workgroup size */
    float forward_ms, backward_ms;
    float default_forward_ms, default_backward_ms;
    int candidates;        /* configurations timed */
} ffdas_das_tuning_t;

/* Autotuner (no reference counterpart; the reference picks its tile width from a fixed
 * heuristic or an environment variable, src/das/das_alg1.cu:209-225).  Times candidate
 * launch parameters of ffdas_das_ex (and, when grad_x != NULL, of ffdas_das_backward) on the
 * handle's stream for exactly this problem signature -- device architecture, storage types,
 * channels, transmits, samples, frames, voxel grid, sparse count, directivity -- and stores
 * the fastest in the library's tuning cache, which every later call with the same signature
 * uses.  `out` and `grad_x` are overwritten (scratch for the timed runs).  Synchronous: it
 * waits for the timed runs.  reps: timed runs per candidate (1..100). */
FFDAS_API ffdas_error_t ffdas_das_autotune(ffdas_handle_t handle, const float *srcpos,
                                           const float *srcdir, float wavenum, float scale,
                                           ffdas_tensor_desc_t x_desc, const void *x,
                                           const float *dstpos, const float *offsets,
                                           const float *weights, int sparse_count,
                                           const int *sparse_indices,
                                           ffdas_tensor_desc_t out_desc, void *out,
                                           ffdas_tensor_desc_t grad_x_desc, void *grad_x,
                                           ffdas_compute_type_t compute_type, ffdas_alg_t alg,
                                           int reps, ffdas_das_tuning_t *result);

/* Reuse of the converted copy of x in the adjoint (no reference counterpart).  Every
 * forward keeps the copy of x it gathers from (FP16/BF16 cast, frame-inner relayout) at the
 * start of the handle's workspace.  While enabled, ffdas_das_backward gathers from that copy
 * instead of making a new one when it was made from the same x pointer, descriptor, storage
 * type and chunking and nothing has used the workspace since; the caller guarantees that x
 * was not modified in between (autograd guarantees it for saved tensors). */
FFDAS_API ffdas_error_t ffdas_set_reuse_x(ffdas_handle_t handle, int enable);

/* Tuning cache persistence: one text line per problem signature. */
FFDAS_API ffdas_error_t ffdas_tune_save(const char *path);
FFDAS_API ffdas_error_t ffdas_tune_load(const char *path);
FFDAS_API ffdas_error_t ffdas_tune_clear(void);
FFDAS_API int ffdas_tune_count(void);

/* Rank truncation filter (truncated-SVD reconstruction): x is an (m, n) matrix per x_desc
 * (R32F, C32F, R64F or C64F; any non-negative element strides); out has the same dims and
 * dtype (dense, row- or column-major, may be x itself).  With p = min(m, n), out keeps the
 * singular components [start, stop) of x counted from the largest, 0 <= start < stop <= p
 * (else FFDAS_ERROR_INVALID_ARGUMENT).  The eigendecomposition of the p x p Gram matrix runs
 * in one workgroup (Jacobi, FFDAS_ERROR_CONVERGENCE_FAILED if 30 sweeps do not converge)
 * for p <= 64, else on rocSOLVER (FFDAS_ERROR_CUSOLVER if it is unavailable or fails); on
 * such a failure out is undefined.  Unlike every other entry point, this one synchronises
 * the handle's stream, once, at its end, to read the solver's status (the reference's also
 * synchronises).
 * Replaces include/ffdas.h:178-186. */
FFDAS_API ffdas_error_t ffdas_truncate_rank(ffdas_handle_t handle, ffdas_tensor_desc_t x_desc,
                                            const void *x, int64_t start, int64_t stop,
                                            ffdas_tensor_desc_t out_desc, void *out);

/* Interpolation from a structured (nz, ny, nx) grid of arbitrary vertex positions to
 * scattered query points.  The plan splits every grid cell into the reference's six
 * tetrahedra and bins them over the grid's bounding box (128 bins per axis); a query is
 * located in the lowest-numbered tetrahedron of its bin whose FP32 barycentric test passes
 * (the reference's result), and evaluates to the vertex of largest weight (NEAREST) or the
 * barycentric combination of the four vertices (LINEAR), or to *fill_value outside the grid.
 * gridpos: device, nz*ny*nx packed float triples, copied into the plan.
 * Replaces include/ffdas.h:248-254. */
FFDAS_API ffdas_error_t ffdas_create_interpolation_plan(ffdas_handle_t handle,
                                                        ffdas_interpolation_plan_t *plan,
                                                        int64_t nx, int64_t ny, int64_t nz,
                                                        const float *gridpos,
                                                        ffdas_interp_mode_t mode);

/* Locates num_querypos points (device, packed float triples) once and keeps the result in
 * the plan; a later ffdas_interpolation with the same num_querypos uses it and ignores its
 * querypos argument (as the reference does).  Replaces include/ffdas.h:256-261. */
FFDAS_API ffdas_error_t ffdas_interpolation_preprocess(ffdas_handle_t handle,
                                                       ffdas_interpolation_plan_t plan,
                                                       int64_t num_querypos,
                                                       const float *querypos);

/* out[b, q] for x ([batch,] nz, ny, nx) per x_desc (R32F, C32F, R64F or C64F, packed) and
 * out (batch, num_querypos) packed, same dtype; fill_value is a host pointer to one element
 * of x's type.  querypos may be NULL when the plan was preprocessed for num_querypos points.
 * Replaces include/ffdas.h:263-272. */
FFDAS_API ffdas_error_t ffdas_interpolation(ffdas_handle_t handle, ffdas_interpolation_plan_t plan,
                                            int64_t num_querypos, const float *querypos,
                                            ffdas_tensor_desc_t x_desc, const void *x, void *out,
                                            const void *fill_value);

/* Releases the plan's device memory (after the handle's stream has finished with it).
 * Replaces include/ffdas.h:274-277. */
FFDAS_API ffdas_error_t ffdas_destroy_interpolation_plan(ffdas_handle_t handle,
                                                         ffdas_interpolation_plan_t plan);

/* Frequency-domain Green's function summation: out[b, t, f] = sum_s x[b, s, f]
 * exp(i k_f r) / r with r = |dstpos[t] - srcpos[s]| (r = rsq * rsqrt(rsq), as the reference;
 * a target on a source gives a non-finite term).  x: (batch, sources, frequencies) C32F or
 * C16F, packed; out: (batch, targets, frequencies) C32F, packed; srcpos, dstpos: device
 * packed float triples; wavenums: device floats, one per frequency.  Arithmetic in FP32
 * (the reference rounds x and the Green's function to FP16 for its tensor-core path).
 * Replaces include/ffdas.h:279-288. */
FFDAS_API ffdas_error_t ffdas_greens(ffdas_handle_t handle, const float *srcpos,
                                     const float *wavenums, ffdas_tensor_desc_t x_desc,
                                     const void *x, const float *dstpos,
                                     ffdas_tensor_desc_t out_desc, void *out);

/* Opt-in kernel timing: while enabled, every kernel the library launches is bracketed by
 * hipEvents on its stream; ffdas_profile_query returns the summed duration [ms] and launch
 * count per kernel name ("das_fwd", "das_fwd_fi", "das_bwd_param", "das_bwd_param_fi",
 * "das_bwd_x", "das_bwd_x_plain", "das_fwd_fi_batched", "relayout_fi", "cast_rows",
 * "transpose_gout", "rank_gram", "rank_gram_reduce", "rank_jacobi", "rank_project",
 * "interp_locate", "interp_frames", "interp_eval", "greens").
 * Counterpart of the reference's NVTX ranges (no reference symbol). */
FFDAS_API ffdas_error_t ffdas_profile_enable(int enable);
FFDAS_API ffdas_error_t ffdas_profile_reset(void);
FFDAS_API ffdas_error_t ffdas_profile_query(const char *name, double *total_ms, int64_t *count);

/* Library version string. */
FFDAS_API const char *ffdas_version(void);

#if defined(__cplusplus)
}
#endif

#endif /* FFDAS_H_ */
