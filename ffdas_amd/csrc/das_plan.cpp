// Host-only argument validation and shape decoding for ffdas_das / ffdas_das_sparse.
//
// Error behaviour follows the reference, in the same order:
//   null descriptors                     -> FFDAS_ERROR_NULL_POINTER   (src/das/das.cu:27-28)
//   unsupported (x, out) dtype pair      -> FFDAS_ERROR_UNSUPPORTED_TYPE (das.cu:40-82)
//   FP64 data with ALG2/ALG4             -> FFDAS_ERROR_UNSUPPORTED_TYPE (das_impl.cuh:37-40)
//   ALG3 / unknown algorithm             -> FFDAS_ERROR_INVALID_ARGUMENT (das_impl.cuh:42-55)
//   sparse with an algorithm other than DEFAULT/ALG1
//                                        -> FFDAS_ERROR_INVALID_ARGUMENT (das_sparse_impl.cuh:38-45)
//   x not 3-D/4-D, out ndim < 1, batch mismatch
//                                        -> FFDAS_ERROR_INVALID_DIMS   (das_utils.h:39-53)
//   sample stride != 1                   -> FFDAS_ERROR_INVALID_ARGUMENT (das_utils.h:66-67)
//   sparse_count <= 0                    -> FFDAS_ERROR_INVALID_ARGUMENT (das_alg1.cu:267-268)
//   COMPUTE_16F on non-float input       -> FFDAS_ERROR_UNSUPPORTED_TYPE (das_alg1.cu:172-178)
// Differences (additive): null geometry pointers are rejected with NULL_POINTER instead of
// faulting; 64-bit frame offsets lift the reference's int32 limit on the whole tensor (only
// one frame of x must stay below 2^31 elements); a non-contiguous output is rejected with
// NON_CONTIGUOUS instead of being written with the wrong layout.
#include <cstring>

#include "internal.h"

namespace ffdas {

namespace {

bool storage_of(ffdas_datatype_t t, Storage &s, bool &cplx) {
    switch (t) {
        case FFDAS_R_16F: s = Storage::F16; cplx = false; return true;
        case FFDAS_R_16BF: s = Storage::BF16; cplx = false; return true;
        case FFDAS_R_32F: s = Storage::F32; cplx = false; return true;
        case FFDAS_R_64F: s = Storage::F64; cplx = false; return true;
        case FFDAS_C_16F: s = Storage::F16; cplx = true; return true;
        case FFDAS_C_16BF: s = Storage::BF16; cplx = true; return true;
        case FFDAS_C_32F: s = Storage::F32; cplx = true; return true;
        case FFDAS_C_64F: s = Storage::F64; cplx = true; return true;
        default: return false;
    }
}

ffdas_datatype_t out_type_for(Storage s, bool cplx) {
    if (s == Storage::F64) return cplx ? FFDAS_C_64F : FFDAS_R_64F;
    return cplx ? FFDAS_C_32F : FFDAS_R_32F;
}

}  // namespace

ffdas_error_t plan_das(const ffdas_tensor_desc &x, const ffdas_tensor_desc &out,
                       const float *srcpos, const float *dstpos, const float *offsets,
                       const float *weights, bool sparse_mode, int sparse_count,
                       const int *sparse_indices, ffdas_compute_type_t ct, ffdas_alg_t alg,
                       DasProblem &p) {
    p = DasProblem();
    bool cplx = false;
    Storage xs;
    if (!storage_of(x.dtype, xs, cplx)) return FFDAS_ERROR_UNSUPPORTED_TYPE;
    if (out.dtype != out_type_for(xs, cplx)) return FFDAS_ERROR_UNSUPPORTED_TYPE;
    p.cplx = cplx;
    p.x_storage = xs;
    p.acc64 = xs == Storage::F64;
    p.sparse = sparse_mode;

    if (xs == Storage::F64 && alg != FFDAS_ALG_DEFAULT && alg != FFDAS_ALG1)
        return FFDAS_ERROR_UNSUPPORTED_TYPE;
    switch (alg) {
        case FFDAS_ALG_DEFAULT: case FFDAS_ALG1: break;
        case FFDAS_ALG2: case FFDAS_ALG4:
            if (p.sparse) return FFDAS_ERROR_INVALID_ARGUMENT;
            break;
        default: return FFDAS_ERROR_INVALID_ARGUMENT;
    }

    const int nd = (int)x.dims.size();
    if (nd != 3 && nd != 4) return FFDAS_ERROR_INVALID_DIMS;
    if (out.dims.empty()) return FFDAS_ERROR_INVALID_DIMS;
    if (x.strides.size() != x.dims.size() || out.strides.size() != out.dims.size())
        return FFDAS_ERROR_INVALID_ARGUMENT;
    p.have_batch = nd == 4;
    const int o = p.have_batch ? 1 : 0;
    for (int64_t d : x.dims)
        if (d <= 0) return FFDAS_ERROR_INVALID_DIMS;
    for (int64_t d : out.dims)
        if (d <= 0) return FFDAS_ERROR_INVALID_DIMS;
    if (x.dims[o + 0] > INT32_MAX || x.dims[o + 1] > INT32_MAX || x.dims[o + 2] > INT32_MAX ||
        x.dims[0] > INT32_MAX)
        return FFDAS_ERROR_INVALID_DIMS;
    p.C = (int)x.dims[o + 0];
    p.T = (int)x.dims[o + 1];
    p.S = (int)x.dims[o + 2];
    p.B = p.have_batch ? (int)x.dims[0] : 1;
    if (p.have_batch && (out.dims.size() < 2 || out.dims[0] != p.B)) return FFDAS_ERROR_INVALID_DIMS;

    p.N = 1;
    p.spatial.clear();
    for (size_t i = o; i < out.dims.size(); i++) {
        if (__builtin_mul_overflow(p.N, out.dims[i], &p.N)) return FFDAS_ERROR_INVALID_DIMS;
        p.spatial.push_back(out.dims[i]);
    }
    if (p.spatial.empty()) p.spatial.push_back(1);
    if (p.N > INT32_MAX) return FFDAS_ERROR_INVALID_DIMS;

    p.xs_c = x.strides[o + 0];
    p.xs_t = x.strides[o + 1];
    p.xs_b = p.have_batch ? x.strides[0] : 0;
    if (x.strides[o + 2] != 1) return FFDAS_ERROR_INVALID_ARGUMENT;
    if (p.xs_c < 0 || p.xs_t < 0 || p.xs_b < 0) return FFDAS_ERROR_INVALID_ARGUMENT;

    // output: spatial part contiguous; batch stride free
    {
        int64_t s = 1;
        for (int i = (int)out.dims.size() - 1; i >= o; i--) {
            if (out.dims[i] > 1 && out.strides[i] != s) return FFDAS_ERROR_NON_CONTIGUOUS;
            s *= out.dims[i];
        }
        p.os_b = p.have_batch ? out.strides[0] : 0;
    }

    if (p.sparse) {
        if (sparse_count <= 0) return FFDAS_ERROR_INVALID_ARGUMENT;
        if (!sparse_indices) return FFDAS_ERROR_NULL_POINTER;
        p.K = sparse_count;
    } else {
        p.K = p.T;
    }

    // storage the kernels gather from
    p.k_storage = xs;
    switch (ct) {
        case FFDAS_COMPUTE_16F:
            if (xs != Storage::F32) return FFDAS_ERROR_UNSUPPORTED_TYPE;
            p.k_storage = Storage::F16;
            break;
        case FFDAS_COMPUTE_16BF:
            if (xs != Storage::F32) return FFDAS_ERROR_UNSUPPORTED_TYPE;
            p.k_storage = Storage::BF16;
            break;
        case FFDAS_COMPUTE_32F:
            // double data gathered from a float copy, accumulated in double
            // (das_alg1.cu:179-184)
            if (xs == Storage::F64) p.k_storage = Storage::F32;
            break;
        case FFDAS_COMPUTE_DEFAULT:
        case FFDAS_COMPUTE_64F:
            break;
        default:
            return FFDAS_ERROR_INVALID_ARGUMENT;
    }
    p.x_needs_copy = p.k_storage != p.x_storage;

    // one frame must be addressable with 32-bit element offsets
    {
        int64_t span = 0;
        span += (int64_t)(p.C - 1) * (p.xs_c < 0 ? -p.xs_c : p.xs_c);
        span += (int64_t)(p.T - 1) * (p.xs_t < 0 ? -p.xs_t : p.xs_t);
        span += p.S;
        if (span >= INT32_MAX) return FFDAS_ERROR_INVALID_DIMS;
    }

    if (!srcpos || !dstpos || !offsets || !weights) return FFDAS_ERROR_NULL_POINTER;
    p.plain = alg == FFDAS_ALG1;
    return FFDAS_SUCCESS;
}

}  // namespace ffdas

extern "C" ffdas_error_t ffdas_das_validate(ffdas_tensor_desc_t x_desc, ffdas_tensor_desc_t out_desc,
                                            const float *srcpos, const float *dstpos,
                                            const float *offsets, const float *weights,
                                            int sparse_count, const int *sparse_indices,
                                            ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    if (!x_desc || !out_desc) return FFDAS_ERROR_NULL_POINTER;
    ffdas::DasProblem p;
    return ffdas::plan_das(*x_desc, *out_desc, srcpos, dstpos, offsets, weights,
                           sparse_indices != nullptr || sparse_count != 0, sparse_count,
                           sparse_indices, compute_type, alg, p);
}
