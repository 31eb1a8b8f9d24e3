// Tuning cache of the delay-and-sum launch parameters (tile shape, channel groups, frames
// per chunk, grad_x workgroup shape), keyed by the problem signature: device architecture,
// storage types, complex data, double accumulation, sparse count, channels, transmits,
// samples, frames, voxel grid and directivity.  Entries are written by ffdas_das_autotune
// (das_api.cpp) and consulted on every call; ffdas_tune_save / ffdas_tune_load keep them
// in a text file, one entry per line:
//   <key> groups fi_frames tile_h tile_w param_groups bwdx_frames bwdx_lg bwdx_threads
#include <cstdio>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>

#include "internal.h"

namespace ffdas {
namespace {

std::mutex g_mu;
std::unordered_map<std::string, DasTune> g_cache;

const char *storage_name(Storage s) {
    switch (s) {
        case Storage::F16: return "f16";
        case Storage::BF16: return "bf16";
        case Storage::F32: return "f32";
        case Storage::F64: return "f64";
    }
    return "?";
}

}  // namespace

std::string tune_key(const ffdas_context &h, const DasProblem &p, bool has_srcdir) {
    std::ostringstream k;
    k << h.arch << "|x=" << storage_name(p.x_storage) << "|k=" << storage_name(p.k_storage)
      << "|c=" << (p.cplx ? 1 : 0) << "|a64=" << (p.acc64 ? 1 : 0) << "|sp="
      << (p.sparse ? p.K : 0) << "|C=" << p.C << "|T=" << p.T << "|S=" << p.S << "|B=" << p.B
      << "|N=";
    for (size_t i = 0; i < p.spatial.size(); i++) k << (i ? "x" : "") << p.spatial[i];
    k << "|dir=" << (has_srcdir ? 1 : 0);
    return k.str();
}

void resolve_tune(const ffdas_context &h, DasProblem &p, bool has_srcdir) {
    if (h.force_tune) {
        p.tune = *h.force_tune;
        return;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(tune_key(h, p, has_srcdir));
    p.tune = it == g_cache.end() ? DasTune() : it->second;
}

void tune_store(const std::string &key, const DasTune &t) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache[key] = t;
}

}  // namespace ffdas

using namespace ffdas;

extern "C" {

ffdas_error_t ffdas_tune_save(const char *path) {
    if (!path) return FFDAS_ERROR_NULL_POINTER;
    std::lock_guard<std::mutex> lk(g_mu);
    std::ofstream f(path);
    if (!f) return FFDAS_ERROR_INVALID_ARGUMENT;
    for (const auto &kv : g_cache) {
        const DasTune &t = kv.second;
        f << kv.first << ' ' << t.groups << ' ' << t.fi_frames << ' ' << t.tile_h << ' '
          << t.tile_w << ' ' << t.param_groups << ' ' << t.bwdx_frames << ' ' << t.bwdx_lg
          << ' ' << t.bwdx_threads << '\n';
    }
    return f ? FFDAS_SUCCESS : FFDAS_ERROR_INVALID_ARGUMENT;
}

ffdas_error_t ffdas_tune_load(const char *path) {
    if (!path) return FFDAS_ERROR_NULL_POINTER;
    std::ifstream f(path);
    if (!f) return FFDAS_ERROR_INVALID_ARGUMENT;
    std::unordered_map<std::string, DasTune> loaded;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream in(line);
        std::string key;
        DasTune t;
        if (!(in >> key >> t.groups >> t.fi_frames >> t.tile_h >> t.tile_w >> t.param_groups >>
              t.bwdx_frames >> t.bwdx_lg >> t.bwdx_threads))
            return FFDAS_ERROR_INVALID_ARGUMENT;
        loaded[key] = t;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto &kv : loaded) g_cache[kv.first] = kv.second;
    return FFDAS_SUCCESS;
}

ffdas_error_t ffdas_tune_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_cache.clear();
    return FFDAS_SUCCESS;
}

int ffdas_tune_count(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return (int)g_cache.size();
}

}  // extern "C"
