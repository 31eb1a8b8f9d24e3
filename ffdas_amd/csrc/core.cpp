// Handle, stream, descriptor, event and error machinery of libffdas.so.
// Semantics follow the reference's core runtime (src/core/context.cu:16-249,
// src/core/tensor.cu:9-42, src/core/type_utils.cu:5-30, src/core/error_utils.cu:3-87)
// re-implemented on the HIP runtime: no BLAS/solver handles, a stream-ordered HIP
// allocation for workspace instead of a caching allocator.
#include <dlfcn.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>

#include "internal.h"

namespace ffdas {

// ---- named ranges and the debug check around C-ABI calls ------------------------------
// The reference brackets every entry point with an NVTX range (src/core/context.cuh:55-77,
// e.g. src/das/das.cu:37) and, when built with FFDAS_DEBUG, synchronises after its launches
// (src/core/error_checking.h:115-126).  Here both are run-time switches: ranges go to roctx
// (librocprofiler-sdk-roctx, loaded on first use; rocprofv3 --marker-trace records them)
// unless FFDAS_ROCTX=0, and FFDAS_DEBUG=1 makes every entry point wait for its stream and
// report an asynchronous failure as FFDAS_ERROR_CUDA_RUNTIME (not while the stream is being
// captured into a graph, where waiting is not allowed).
namespace {
struct Roctx {
    int (*push)(const char *) = nullptr;
    int (*pop)() = nullptr;
};

const Roctx &roctx() {
    static Roctx r;
    static std::once_flag once;
    std::call_once(once, [] {
        const char *off = std::getenv("FFDAS_ROCTX");
        if (off && std::strcmp(off, "0") == 0) return;
        void *lib = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!lib) lib = dlopen("librocprofiler-sdk-roctx.so", RTLD_NOW | RTLD_LOCAL);
        if (!lib) return;
        auto push = reinterpret_cast<int (*)(const char *)>(dlsym(lib, "roctxRangePushA"));
        auto pop = reinterpret_cast<int (*)()>(dlsym(lib, "roctxRangePop"));
        if (push && pop) {
            r.push = push;
            r.pop = pop;
        }
    });
    return r;
}
}  // namespace

bool debug_sync_enabled() {
    const char *e = std::getenv("FFDAS_DEBUG");
    return e && *e && std::strcmp(e, "0") != 0;
}

CallScope::CallScope(const ffdas_context *h, const char *name) : h_(h) {
    const Roctx &r = roctx();
    if (r.push) {
        r.push(name);
        pushed_ = true;
    }
}

CallScope::~CallScope() {
    if (pushed_) roctx().pop();
}

ffdas_error_t CallScope::done(ffdas_error_t status) const {
    if (status != FFDAS_SUCCESS || !h_ || !debug_sync_enabled()) return status;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h_->stream, &cs) != hipSuccess) return FFDAS_ERROR_CUDA_RUNTIME;
    if (cs != hipStreamCaptureStatusNone) return status;
    if (hipStreamSynchronize(h_->stream) != hipSuccess) return FFDAS_ERROR_CUDA_RUNTIME;
    if (hipGetLastError() != hipSuccess) return FFDAS_ERROR_CUDA_RUNTIME;
    return status;
}

ffdas_error_t hip_status(hipError_t e, ffdas_error_t kind) {
    if (e == hipSuccess) return FFDAS_SUCCESS;
    if (e == hipErrorOutOfMemory) return FFDAS_ERROR_DEVICE_MEMORY_INSUFFICIENT;
    if (e == hipErrorNoDevice) return FFDAS_ERROR_DEVICE_NOT_AVAILABLE;
    if (e == hipErrorInvalidDevice) return FFDAS_ERROR_INVALID_DEVICE;
    return kind;
}

}  // namespace ffdas

using ffdas::hip_status;

ffdas_error_t ffdas_context::check_device() const {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return FFDAS_ERROR_CUDA_RUNTIME;
    return cur == device_id ? FFDAS_SUCCESS : FFDAS_ERROR_DEVICE_MISMATCH;
}

ffdas_error_t ffdas_context::workspace(size_t bytes, void **ptr) {
    if (bytes == 0) {
        *ptr = nullptr;
        return FFDAS_SUCCESS;
    }
    // Under stream capture nothing may allocate or synchronise: the workspace must already
    // be large enough (one call of the same shape before capturing).
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cs) != hipSuccess) cs = hipStreamCaptureStatusNone;
    const bool capturing = cs != hipStreamCaptureStatusNone;
    if (capturing) {
        if (bytes > ws_bytes) return FFDAS_ERROR_ALLOCATION_FAILED;
        *ptr = ws;
        return FFDAS_SUCCESS;
    }
    if (ws && ws_stream != stream) {
        // the previous owner stream may still read the workspace
        if (hipStreamSynchronize(ws_stream) != hipSuccess) return FFDAS_ERROR_CUDA_SYNC_FAILED;
        ws_stream = stream;
    }
    if (bytes > ws_bytes) {
        if (ws) {
            retired.push_back(ws);
            ws = nullptr;
            ws_bytes = 0;
        }
        size_t want = bytes + bytes / 4;  // headroom against repeated growth
        want = (want + 4095) & ~size_t(4095);
        hipError_t e = hipMallocAsync(&ws, want, stream);
        if (e != hipSuccess) {
            ws = nullptr;
            return e == hipErrorOutOfMemory ? FFDAS_ERROR_DEVICE_MEMORY_INSUFFICIENT
                                            : FFDAS_ERROR_ALLOCATION_FAILED;
        }
        ws_bytes = want;
        ws_stream = stream;
    }
    *ptr = ws;
    return FFDAS_SUCCESS;
}

ffdas_error_t ffdas_context::side_stream(hipStream_t *s) {
    if (!side) {
        hipStream_t st = nullptr;
        hipEvent_t a = nullptr, b = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&a, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b, hipEventDisableTiming);
        if (e != hipSuccess) {
            if (st) (void)hipStreamDestroy(st);
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
            return hip_status(e);
        }
        side = st;
        fork_ev = a;
        join_ev = b;
    }
    *s = side;
    return FFDAS_SUCCESS;
}

extern "C" {

const char *ffdas_version(void) { return "ffdas-mi355x 0.1.0 (gfx950)"; }

size_t ffdas_type_size(ffdas_datatype_t t) {
    switch (t) {
        case FFDAS_R_16F: case FFDAS_R_16BF: case FFDAS_R_16I: return 2;
        case FFDAS_C_16F: case FFDAS_C_16BF: case FFDAS_C_16I: return 4;
        case FFDAS_R_32F: case FFDAS_R_32I: return 4;
        case FFDAS_C_32F: case FFDAS_C_32I: return 8;
        case FFDAS_R_64F: return 8;
        case FFDAS_C_64F: return 16;
        default: return 0;
    }
}

const char *ffdas_error_string(ffdas_error_t e) {
    switch (e) {
        case FFDAS_SUCCESS: return "Success";
        case FFDAS_ERROR_INVALID_ARGUMENT: return "Invalid argument";
        case FFDAS_ERROR_INVALID_DIMS: return "Invalid dimensions";
        case FFDAS_ERROR_UNSUPPORTED_TYPE: return "Unsupported data type";
        case FFDAS_ERROR_NULL_POINTER: return "Null pointer argument";
        case FFDAS_ERROR_INVALID_HANDLE: return "Invalid handle";
        case FFDAS_ERROR_NON_CONTIGUOUS: return "Tensor is not contiguous";
        case FFDAS_ERROR_DIMS_TOO_LARGE: return "Dimensions too large";
        case FFDAS_ERROR_ALLOCATION_FAILED: return "Memory allocation failed";
        case FFDAS_ERROR_DEVICE_MEMORY_INSUFFICIENT: return "Insufficient device memory";
        case FFDAS_ERROR_HOST_MEMORY_INSUFFICIENT: return "Insufficient host memory";
        case FFDAS_ERROR_INVALID_DEVICE: return "Invalid device";
        case FFDAS_ERROR_DEVICE_NOT_SUPPORTED: return "Device not supported (gfx950 required)";
        case FFDAS_ERROR_INSUFFICIENT_COMPUTE_CAPABILITY: return "Insufficient compute capability";
        case FFDAS_ERROR_DEVICE_NOT_AVAILABLE: return "Device not available";
        case FFDAS_ERROR_DEVICE_MISMATCH: return "Device mismatch";
        case FFDAS_ERROR_CUDA_DRIVER: return "GPU driver error";
        case FFDAS_ERROR_CUDA_RUNTIME: return "HIP runtime error";
        case FFDAS_ERROR_CUDA_LAUNCH_FAILED: return "HIP kernel launch failed";
        case FFDAS_ERROR_CUDA_SYNC_FAILED: return "HIP synchronization failed";
        case FFDAS_ERROR_CUDA_MEMCPY_FAILED: return "HIP memory copy failed";
        case FFDAS_ERROR_CUBLAS: return "BLAS library error";
        case FFDAS_ERROR_CUSOLVER: return "Solver library error";
        case FFDAS_ERROR_CUFFT: return "FFT library error";
        case FFDAS_ERROR_NVTX: return "Tracing error";
        case FFDAS_ERROR_ALGORITHM_FAILED: return "Algorithm execution failed";
        case FFDAS_ERROR_CONVERGENCE_FAILED: return "Algorithm failed to converge";
        case FFDAS_ERROR_NUMERICAL_INSTABILITY: return "Numerical instability detected";
        case FFDAS_ERROR_NOT_INITIALIZED: return "Library not initialized";
        case FFDAS_ERROR_ALREADY_INITIALIZED: return "Library already initialized";
        case FFDAS_ERROR_INVALID_STATE: return "Invalid library state";
        case FFDAS_ERROR_OUT_OF_BOUNDS_INDEX: return "Out of bounds index";
        case FFDAS_ERROR_INTERNAL: return "Internal error";
        case FFDAS_ERROR_UNKNOWN: return "Unknown error";
        default: return "Unrecognized error code";
    }
}

ffdas_error_t ffdas_create_tensor_desc(ffdas_tensor_desc_t *desc, int64_t ndim,
                                       const int64_t *dims, const int64_t *strides,
                                       ffdas_datatype_t dtype) {
    if (!desc || ndim <= 0 || !dims) return FFDAS_ERROR_INVALID_ARGUMENT;
    if (ffdas_type_size(dtype) == 0) return FFDAS_ERROR_UNSUPPORTED_TYPE;
    ffdas_tensor_desc *d = new (std::nothrow) ffdas_tensor_desc();
    if (!d) return FFDAS_ERROR_ALLOCATION_FAILED;
    try {
        d->dims.assign(dims, dims + ndim);
        d->strides.resize(ndim);
        if (strides) {
            d->strides.assign(strides, strides + ndim);
        } else {
            int64_t s = 1;
            for (int64_t i = ndim - 1; i >= 0; i--) {
                d->strides[i] = s;
                // a stride that does not fit in 64 bits describes no addressable tensor
                if (i > 0 && __builtin_mul_overflow(s, dims[i], &s)) {
                    delete d;
                    return FFDAS_ERROR_DIMS_TOO_LARGE;
                }
            }
        }
    } catch (...) {
        delete d;
        return FFDAS_ERROR_ALLOCATION_FAILED;
    }
    d->dtype = dtype;
    *desc = d;
    return FFDAS_SUCCESS;
}

ffdas_error_t ffdas_destroy_tensor_desc(ffdas_tensor_desc_t desc) {
    delete desc;
    return FFDAS_SUCCESS;
}

ffdas_error_t ffdas_create(ffdas_handle_t *handle) {
    if (!handle) return FFDAS_ERROR_NULL_POINTER;
    *handle = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return FFDAS_ERROR_DEVICE_NOT_AVAILABLE;
    ffdas_context *h = new (std::nothrow) ffdas_context();
    if (!h) return FFDAS_ERROR_HOST_MEMORY_INSUFFICIENT;
    ffdas_error_t err = FFDAS_SUCCESS;
    hipDeviceProp_t prop;
    if ((e = hipGetDevice(&h->device_id)) != hipSuccess) {
        err = hip_status(e);
    } else if ((e = hipGetDeviceProperties(&prop, h->device_id)) != hipSuccess) {
        err = hip_status(e);
    } else {
        h->arch = prop.gcnArchName;
        h->cu_count = prop.multiProcessorCount;
        h->lds_bytes = (int)prop.maxSharedMemoryPerMultiProcessor;
        if (h->arch.compare(0, 6, "gfx950") != 0) err = FFDAS_ERROR_DEVICE_NOT_SUPPORTED;
    }
    if (err != FFDAS_SUCCESS) {
        delete h;
        return err;
    }
    *handle = h;
    return FFDAS_SUCCESS;
}

ffdas_error_t ffdas_destroy(ffdas_handle_t h) {
    if (!h) return FFDAS_ERROR_INVALID_HANDLE;
    ffdas_error_t err = FFDAS_SUCCESS;
    ffdas::release_rank_solver(*h);
    if (h->status_host) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipHostFree(h->status_host);
    }
    if (h->side) {
        (void)hipStreamSynchronize(h->side);
        (void)hipStreamDestroy(h->side);
        (void)hipEventDestroy(h->fork_ev);
        (void)hipEventDestroy(h->join_ev);
    }
    if (h->ws) h->retired.push_back(h->ws);
    for (void *p : h->retired) {
        // release in stream order; if the runtime is already torn down, just leak
        hipError_t e = hipFreeAsync(p, h->ws_stream);
        if (e != hipSuccess && e != hipErrorDeinitialized && e != hipErrorContextIsDestroyed)
            err = FFDAS_ERROR_CUDA_RUNTIME;
    }
    delete h;
    return err;
}

ffdas_error_t ffdas_set_stream(ffdas_handle_t h, uintptr_t stream) {
    if (!h) return FFDAS_ERROR_INVALID_HANDLE;
    h->stream = reinterpret_cast<hipStream_t>(stream);
    return FFDAS_SUCCESS;
}

ffdas_error_t ffdas_get_stream(ffdas_handle_t h, uintptr_t *stream) {
    if (!h) return FFDAS_ERROR_INVALID_HANDLE;
    if (!stream) return FFDAS_ERROR_NULL_POINTER;
    *stream = reinterpret_cast<uintptr_t>(h->stream);
    return FFDAS_SUCCESS;
}

ffdas_error_t ffdas_device_count(int *count) {
    if (!count) return FFDAS_ERROR_NULL_POINTER;
    hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) {
        *count = 0;
        return FFDAS_SUCCESS;
    }
    return hip_status(e);
}

ffdas_error_t ffdas_device_get(int *device) {
    if (!device) return FFDAS_ERROR_NULL_POINTER;
    return hip_status(hipGetDevice(device));
}

ffdas_error_t ffdas_device_set(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return hip_status(e);
    if (device < 0 || device >= count) return FFDAS_ERROR_INVALID_DEVICE;
    return hip_status(hipSetDevice(device));
}

ffdas_error_t ffdas_event_create(uintptr_t *event) {
    if (!event) return FFDAS_ERROR_NULL_POINTER;
    hipEvent_t ev;
    hipError_t e = hipEventCreate(&ev);
    if (e != hipSuccess) return hip_status(e);
    *event = reinterpret_cast<uintptr_t>(ev);
    return FFDAS_SUCCESS;
}

ffdas_error_t ffdas_event_destroy(uintptr_t event) {
    if (event == 0) return FFDAS_SUCCESS;
    return hip_status(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
}

ffdas_error_t ffdas_event_record(ffdas_handle_t h, uintptr_t event) {
    if (!h) return FFDAS_ERROR_INVALID_HANDLE;
    if (event == 0) return FFDAS_ERROR_INVALID_ARGUMENT;
    FFDAS_TRY(h->check_device());
    return hip_status(hipEventRecord(reinterpret_cast<hipEvent_t>(event), h->stream));
}

ffdas_error_t ffdas_event_synchronize(uintptr_t event) {
    if (event == 0) return FFDAS_ERROR_INVALID_ARGUMENT;
    return hip_status(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)));
}

ffdas_error_t ffdas_event_elapsed_time(uintptr_t start, uintptr_t stop, float *ms) {
    if (!ms) return FFDAS_ERROR_NULL_POINTER;
    if (start == 0 || stop == 0) return FFDAS_ERROR_INVALID_ARGUMENT;
    return hip_status(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                          reinterpret_cast<hipEvent_t>(stop)));
}

}  // extern "C"
