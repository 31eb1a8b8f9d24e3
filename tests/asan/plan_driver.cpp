// Host AddressSanitizer / UndefinedBehaviorSanitizer run of the C-ABI's host-side argument
// handling (ffdas_amd/csrc/core.cpp descriptors and error strings, das_plan.cpp validation via
// ffdas_das_validate): descriptors of every rank 1-6, zero-sized, huge, negative and odd
// strides, every element and compute type, every algorithm, sparse counts and null pointers.
// No device is touched.  Exit status 0 and no sanitizer report is the pass condition.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "ffdas.h"
#include "internal.h"

// core.cpp's ffdas_destroy releases the rank-truncation solver (rank_api.cpp, which needs the
// device libraries); no handle is created here, so it is never called.
namespace ffdas {
void release_rank_solver(ffdas_context &) {}
}  // namespace ffdas

static ffdas_tensor_desc_t desc(std::vector<int64_t> d, std::vector<int64_t> s,
                                ffdas_datatype_t t) {
    ffdas_tensor_desc_t h = nullptr;
    if (ffdas_create_tensor_desc(&h, (int64_t)d.size(), d.data(), s.empty() ? nullptr : s.data(),
                                 t) != FFDAS_SUCCESS)
        return nullptr;
    return h;
}

int main() {
    const ffdas_datatype_t types[] = {FFDAS_R_16F, FFDAS_C_16F, FFDAS_R_32F, FFDAS_C_32F,
                                      FFDAS_R_64F, FFDAS_C_64F, FFDAS_R_32I, FFDAS_C_32I,
                                      FFDAS_R_16I, FFDAS_C_16I, FFDAS_R_16BF, FFDAS_C_16BF};
    const std::vector<std::vector<int64_t>> xdims = {
        {4}, {3, 8}, {5, 3, 64}, {2, 5, 3, 64}, {2, 5, 3, 64, 1}, {1, 1, 1, 1, 1, 1},
        {0, 5, 3, 64}, {2, 5, 0, 64}, {2, 5, 3, 1}, {1 << 20, 1 << 20, 3, 64}, {-1, 5, 3, 64}};
    const std::vector<std::vector<int64_t>> odims = {
        {4}, {7, 9}, {2, 7, 9}, {2, 63}, {0, 7}, {1 << 30, 1 << 30, 9}, {3, 7, 9, 1, 1}};
    float geo[16] = {0};
    int idx[16] = {0};
    long checked = 0;
    for (auto xt : types)
        for (auto ot : types)
            for (const auto &xd : xdims)
                for (const auto &od : odims)
                    for (int strided = 0; strided < 3; strided++) {
                        std::vector<int64_t> xs;
                        if (strided == 1) {        // padded rows, negative channel stride
                            xs.assign(xd.size(), 1);
                            int64_t st = 1;
                            for (size_t i = xd.size(); i-- > 0;) {
                                xs[i] = (i == 0 && xd.size() > 1) ? -st : st;
                                st *= (xd[i] > 0 ? xd[i] : 1) + 3;
                            }
                        } else if (strided == 2) { // non-unit sample stride
                            xs.assign(xd.size(), 2);
                        }
                        ffdas_tensor_desc_t x = desc(xd, xs, xt), o = desc(od, {}, ot);
                        for (int ct = 0; ct <= 5; ct++)
                            for (int alg = 0; alg <= 5; alg++)
                                for (int sc : {0, 3, -2}) {
                                    const bool sp = sc != 0;
                                    (void)ffdas_das_validate(x, o, geo, geo, geo, geo, sc,
                                                             sp ? idx : nullptr,
                                                             (ffdas_compute_type_t)ct,
                                                             (ffdas_alg_t)alg);
                                    (void)ffdas_das_validate(x, o, nullptr, geo, nullptr, geo, sc,
                                                             nullptr, (ffdas_compute_type_t)ct,
                                                             (ffdas_alg_t)alg);
                                    checked += 2;
                                }
                        ffdas_destroy_tensor_desc(x);
                        ffdas_destroy_tensor_desc(o);
                    }
    for (int e = -1; e <= 102; e++) (void)ffdas_error_string((ffdas_error_t)e);
    for (int t = -1; t <= 20; t++) (void)ffdas_type_size((ffdas_datatype_t)t);
    ffdas_tensor_desc_t bad = nullptr;
    (void)ffdas_create_tensor_desc(&bad, 0, nullptr, nullptr, FFDAS_R_32F);
    (void)ffdas_create_tensor_desc(nullptr, 1, nullptr, nullptr, FFDAS_R_32F);
    std::printf("plan sanitizer run: %ld validations clean\n", checked);
    return 0;
}
