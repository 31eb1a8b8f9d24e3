// Host AddressSanitizer / UndefinedBehaviorSanitizer run of the CPU oracle (oracle/das_oracle.cpp):
// forward and adjoint, FP64 and FP32, real and complex, dense and sparse, on small random
// problems whose delays fall inside, at the edges of and outside the sample range, with
// sparse indices outside [0, T), directivity, a non-unit scale and beta accumulation.
// Exit status 0 and no sanitizer report is the pass condition.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

extern "C" {
int ffdas_oracle_das(int real_bits, int nthreads, int B, int C, int T, int S, int N, int K,
                     int is_complex, const float *srcpos, const float *srcdir, float wavenum,
                     double scale, const void *x, const float *dstpos, const float *offsets,
                     const float *weights, const int *sparse, double beta_re, double beta_im,
                     void *out);
int ffdas_oracle_das_adjoint(int real_bits, int nthreads, int B, int C, int T, int S, int N, int K,
                             int is_complex, const float *srcpos, const float *srcdir,
                             float wavenum, double scale, const void *x, const float *dstpos,
                             const float *offsets, const float *weights, const int *sparse,
                             const void *gout, void *gx, void *goffsets, void *gweights,
                             void *gscale);
}

template <class R>
int run(std::mt19937 &rng, int bits, bool cplx, bool sparse, bool dir) {
    const int B = 2, C = 5, T = 3, S = 40, N = 23, K = sparse ? 4 : T;
    const int cw = cplx ? 2 : 1;
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    std::vector<float> src(3 * C), sd(4 * C), dst(3 * N), ofs(K * N), wts(K * N);
    for (auto &v : src) v = 5.0f * u(rng);
    for (int m = 0; m < C; m++) {
        sd[4 * m] = 0.0f; sd[4 * m + 1] = 0.0f; sd[4 * m + 2] = 1.0f; sd[4 * m + 3] = 0.3f;
    }
    for (int n = 0; n < N; n++) {
        dst[3 * n] = 8.0f * u(rng); dst[3 * n + 1] = 2.0f * u(rng);
        dst[3 * n + 2] = 10.0f + 8.0f * u(rng);
    }
    // delays from well below 0 to beyond S - 1, and exactly at the last valid sample
    for (int i = 0; i < K * N; i++) ofs[i] = 25.0f * u(rng) + (i % 7 == 0 ? 30.0f : 0.0f);
    ofs[0] = (float)(S - 2) - 12.0f;
    for (auto &v : wts) v = u(rng);
    std::vector<int> sp(K * N);
    std::uniform_int_distribution<int> ui(-1, T);           // includes -1 and T: outside [0, T)
    for (auto &v : sp) v = ui(rng);
    std::vector<R> x((size_t)B * C * T * S * cw), out((size_t)B * N * cw), g((size_t)B * N * cw);
    for (auto &v : x) v = (R)u(rng);
    for (auto &v : g) v = (R)u(rng);
    for (auto &v : out) v = (R)u(rng);
    std::vector<R> gx(x.size()), go(K * N), gw(K * N), gs(1);
    int e = ffdas_oracle_das(bits, 3, B, C, T, S, N, K, cplx, src.data(), dir ? sd.data() : nullptr,
                             0.7f, 1.013, x.data(), dst.data(), ofs.data(), wts.data(),
                             sparse ? sp.data() : nullptr, 0.5, cplx ? 0.25 : 0.0, out.data());
    if (e) return e;
    e = ffdas_oracle_das_adjoint(bits, 3, B, C, T, S, N, K, cplx, src.data(),
                                 dir ? sd.data() : nullptr, 0.7f, 0.987, x.data(), dst.data(),
                                 ofs.data(), wts.data(), sparse ? sp.data() : nullptr, g.data(),
                                 gx.data(), go.data(), gw.data(), gs.data());
    return e;
}

int main() {
    std::mt19937 rng(7);
    int cases = 0;
    for (int bits : {64, 32})
        for (int cplx = 0; cplx < 2; cplx++)
            for (int sparse = 0; sparse < 2; sparse++)
                for (int dir = 0; dir < 2; dir++) {
                    const int e = bits == 64 ? run<double>(rng, bits, cplx, sparse, dir)
                                             : run<float>(rng, bits, cplx, sparse, dir);
                    if (e) {
                        std::fprintf(stderr, "oracle rejected case %d: code %d\n", cases, e);
                        return 1;
                    }
                    cases++;
                }
    std::printf("oracle sanitizer run: %d cases clean\n", cases);
    return 0;
}
