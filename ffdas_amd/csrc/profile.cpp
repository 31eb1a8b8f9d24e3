// Per-kernel GPU timing (opt-in): when enabled, every kernel launch of the library is
// bracketed by a pair of hipEvents recorded on the launch stream; totals per kernel name are
// resolved on query.  The counterpart of the reference's NVTX ranges around its launches
// (src/core/context.cuh:55-77), with durations instead of trace markers.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"

namespace ffdas {
namespace {

struct Pending {
    std::string name;
    hipEvent_t a, b;
};

struct Totals {
    double ms = 0.0;
    long count = 0;
};

std::mutex g_mu;
bool g_enabled = false;
std::vector<Pending> g_pending;
std::map<std::string, Totals> g_totals;

void resolve_locked() {
    for (auto &p : g_pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto &t = g_totals[p.name];
            t.ms += ms;
            t.count += 1;
        }
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    g_pending.clear();
}

}  // namespace

bool profiling_enabled() { return g_enabled; }

void *profile_begin(hipStream_t st) {
    if (!g_enabled) return nullptr;
    hipEvent_t a;
    if (hipEventCreate(&a) != hipSuccess) return nullptr;
    hipEventRecord(a, st);
    return a;
}

void profile_end(void *begin, hipStream_t st, const char *name) {
    if (!begin) return;
    hipEvent_t b;
    if (hipEventCreate(&b) != hipSuccess) {
        hipEventDestroy(static_cast<hipEvent_t>(begin));
        return;
    }
    hipEventRecord(b, st);
    std::lock_guard<std::mutex> lk(g_mu);
    g_pending.push_back({name, static_cast<hipEvent_t>(begin), b});
    if (g_pending.size() > 4096) resolve_locked();
}

}  // namespace ffdas

extern "C" {

ffdas_error_t ffdas_profile_enable(int enable) {
    std::lock_guard<std::mutex> lk(ffdas::g_mu);
    ffdas::g_enabled = enable != 0;
    return FFDAS_SUCCESS;
}

ffdas_error_t ffdas_profile_reset(void) {
    std::lock_guard<std::mutex> lk(ffdas::g_mu);
    ffdas::resolve_locked();
    ffdas::g_totals.clear();
    return FFDAS_SUCCESS;
}

ffdas_error_t ffdas_profile_query(const char *name, double *total_ms, int64_t *count) {
    if (!name || !total_ms || !count) return FFDAS_ERROR_NULL_POINTER;
    std::lock_guard<std::mutex> lk(ffdas::g_mu);
    ffdas::resolve_locked();
    auto it = ffdas::g_totals.find(name);
    *total_ms = it == ffdas::g_totals.end() ? 0.0 : it->second.ms;
    *count = it == ffdas::g_totals.end() ? 0 : it->second.count;
    return FFDAS_SUCCESS;
}

}  // extern "C"
