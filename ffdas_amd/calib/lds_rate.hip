// LDS atomic throughput probe for gfx950 (a measurement tool, not part of libffdas.so).
//
// The grad_x adjoint of delay-and-sum accumulates into LDS with 64-bit integer atomics, so
// its on-chip ceiling is the LDS atomic rate.  This probe measures that rate on the device
// it runs on, for a given per-lane address pattern: every lane of every wave adds into the
// LDS word pattern[lane] (plus a fixed multiple of 64 words per unrolled step, which keeps
// the bank of each lane), 16 atomics per loop iteration, no returned values; a lane whose
// pattern entry is -1 takes part in no atomic (partly active instructions).  The caller
// converts the time into lane-ops per clock per CU.
//
//   op 0: ds_add_u64 on 64-bit words      op 1: ds_add_u32 on 32-bit words
//   op 2: ds_add_f32 on 32-bit words
//
// Built by ffdas_amd/build.py into ffdas_amd/calib/libffdas_calib.so; used by bench.py and
// ffdas_amd/utils/calib.py.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kUnroll = 16;
constexpr int kStrideWords = 64;       // per unrolled step: 512 B (u64) / 256 B (u32)
constexpr int kLdsBytes = 32 * 1024;

template <int OP>
__global__ __launch_bounds__(512) void lds_atomic_probe(const int *__restrict__ pattern, int reps,
                                                       unsigned long long *__restrict__ sink) {
    __shared__ __align__(16) unsigned long long lds[kLdsBytes / 8];
    for (int i = threadIdx.x; i < kLdsBytes / 8; i += blockDim.x) lds[i] = 0ull;
    __syncthreads();
    const int off = pattern[threadIdx.x & 63];
    if (off >= 0)
    for (int r = 0; r < reps; r++) {
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            if constexpr (OP == 0)
                atomicAdd(lds + off + u * kStrideWords, 1ull);
            else if constexpr (OP == 1)
                atomicAdd(reinterpret_cast<unsigned *>(lds) + off + u * kStrideWords, 1u);
            else
                atomicAdd(reinterpret_cast<float *>(lds) + off + u * kStrideWords, 1.0f);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) sink[blockIdx.x] = lds[0] + lds[kLdsBytes / 8 - 1];
}

}  // namespace

// Times one launch of `blocks` workgroups of `threads` threads, each lane doing reps * 16
// atomics at pattern[lane] (64 host ints, word offsets, -1 for an idle lane; the largest
// must stay below (32 KiB / word size) - 15 * 64).  Writes the kernel time [ms] and returns a hipError_t
// value (0 on success).  Synchronous; allocates and frees its own buffers.
extern "C" __attribute__((visibility("default"))) int
ffdas_calib_lds_atomics(int op, const int *pattern, int reps, int blocks, int threads,
                        float *ms_out) {
    if (!pattern || !ms_out || op < 0 || op > 2 || reps < 1 || blocks < 1 || threads < 64 ||
        threads > 512 || threads % 64)
        return (int)hipErrorInvalidValue;
    const int words = op == 0 ? kLdsBytes / 8 : kLdsBytes / 4;
    for (int i = 0; i < 64; i++)
        if (pattern[i] < -1 || pattern[i] + (kUnroll - 1) * kStrideWords >= words)
            return (int)hipErrorInvalidValue;
    int *dpat = nullptr;
    unsigned long long *sink = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipMalloc(&dpat, 64 * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&sink, (size_t)blocks * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemcpy(dpat, pattern, 64 * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    auto launch = [&]() {
        if (op == 0) lds_atomic_probe<0><<<blocks, threads>>>(dpat, reps, sink);
        else if (op == 1) lds_atomic_probe<1><<<blocks, threads>>>(dpat, reps, sink);
        else lds_atomic_probe<2><<<blocks, threads>>>(dpat, reps, sink);
        return hipGetLastError();
    };
    if (e == hipSuccess) e = launch();                    // warm-up
    if (e == hipSuccess) e = hipEventRecord(a, nullptr);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(b, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms_out, a, b);
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    if (dpat) hipFree(dpat);
    if (sink) hipFree(sink);
    return (int)e;
}

// Compute units and peak engine clock [kHz] of the current device (hipDeviceProp_t).
extern "C" __attribute__((visibility("default"))) int
ffdas_calib_device_info(int *cus, int *clock_khz) {
    if (!cus || !clock_khz) return (int)hipErrorInvalidValue;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(clock_khz, hipDeviceAttributeClockRate, dev);
    return (int)e;
}
