// ffdas_das_noalloc: the delay-and-sum forward of ffdas_das for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  No counterpart in the reference.
//
// The only device memory ffdas_das needs beyond its inputs and outputs is the converted copy
// of the channel data the kernels gather from (the FP16/BF16 cast and the frame-inner
// relayout).  Here that copy lives in `scratch`, a file-scope device array of kScratchBytes
// (256 MiB) that every GPU the library runs on holds until the process ends.  Its address on
// a device is looked up once, with hipGetSymbolAddress, at the first call on that device, and
// kept per device; the host-side name `scratch` is never used as a device address.  A problem
// whose copy exceeds kScratchBytes returns FFDAS_ERROR_DIMS_TOO_LARGE; nothing in this path
// calls hipMalloc, hipFree or an allocating library routine.  The copy is written completely
// by each call (padding frames included), so nothing has to be cleared between calls.
//
// All calls on one GPU share that GPU's array: calls that overlap on one GPU (on different
// streams or host threads without ordering) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

namespace {

constexpr size_t kScratchBytes = size_t(256) << 20;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be current.
ffdas_error_t scratch_on(int device, void **ptr) {
    if (device < 0 || device >= kMaxDevices) return FFDAS_ERROR_INVALID_DEVICE;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return ffdas::hip_status(e, FFDAS_ERROR_CUDA_RUNTIME);
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return FFDAS_SUCCESS;
}

ffdas_error_t das_noalloc_call(ffdas_handle_t handle, const float *srcpos, const float *srcdir,
                               float wavenum, ffdas_tensor_desc_t x_desc, const void *x,
                               const float *dstpos, const float *offsets, const float *weights,
                               const void *beta, ffdas_tensor_desc_t out_desc, void *out,
                               ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    if (!handle) return FFDAS_ERROR_INVALID_HANDLE;
    FFDAS_TRY(handle->check_device());
    const int device = handle->device_id;
    const ffdas::WorkspaceFn ws = [device](size_t bytes, void **ptr) -> ffdas_error_t {
        if (bytes == 0) {
            *ptr = nullptr;
            return FFDAS_SUCCESS;
        }
        if (bytes > kScratchBytes) return FFDAS_ERROR_DIMS_TOO_LARGE;
        return scratch_on(device, ptr);
    };
    // the copy is not recorded for reuse by ffdas_das_backward, which reads the handle's own
    // workspace
    return ffdas::das_forward_with(handle, srcpos, srcdir, wavenum, 1.0f, x_desc, x, dstpos,
                                   offsets, weights, false, 0, nullptr, beta, out_desc, out,
                                   compute_type, alg, ws, false);
}

}  // namespace

extern "C" ffdas_error_t ffdas_das_noalloc(ffdas_handle_t handle, const float *srcpos,
                                           const float *srcdir, float wavenum,
                                           ffdas_tensor_desc_t x_desc, const void *x,
                                           const float *dstpos, const float *offsets,
                                           const float *weights, const void *beta,
                                           ffdas_tensor_desc_t out_desc, void *out,
                                           ffdas_compute_type_t compute_type, ffdas_alg_t alg) {
    ffdas::CallScope cs(handle, "ffdas_das_noalloc");
    return cs.done(das_noalloc_call(handle, srcpos, srcdir, wavenum, x_desc, x, dstpos, offsets,
                                    weights, beta, out_desc, out, compute_type, alg));
}
