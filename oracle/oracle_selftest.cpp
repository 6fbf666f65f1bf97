// oracle_selftest.cpp -- drives every entry point of the CPU oracle (rd_oracle.cpp) on seeded
// clouds, for the sanitizer build (`make -C oracle sanitize`: -fsanitize=address,undefined,
// aborting on the first report).  TEST INFRASTRUCTURE ONLY; tests/test_oracle_sanitize.py runs
// it.  Exit status 0 = every call returned and the results pass a few sanity checks.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

extern "C" {
#define DECL(SUF, T)                                                                                              \
    long oracle_choose_##SUF(const T*, long, int, T, T*, long*);                                                \
    long oracle_evolve_iter_##SUF(const T*, long, T*, long, int, T, int*, int*, T*, unsigned char*);            \
    long oracle_evolve_##SUF(const T*, long, T*, long, int, T, long);                                           \
    long oracle_evolve_omp_iter_##SUF(const T*, long, T*, long, int, T, int*, T*);                              \
    long oracle_decimate_##SUF(T*, long, int, T, int, long*);                                                   \
    long oracle_rd_##SUF(const T*, long, int, T, T, long, T*, long*);                                           \
    long oracle_rd_own_##SUF(const T*, long, long, int, T, T, long, T*, long*);                                 \
    long oracle_tiled_##SUF(int, const T*, long, int, T, T, long, const long*, int, long, T*, long);            \
    long oracle_evolve_tape_##SUF(const T*, long, T*, long, int, T, long, int*, int*, unsigned char*);          \
    long oracle_evolve_replay_##SUF(const T*, long, T*, long, int, long, const int*, const unsigned char*);     \
    long oracle_evolve_vjp_##SUF(long, long, int, long, const int*, const int*, const unsigned char*, T*, T*);  \
    long oracle_set_distances_##SUF(const T*, long, const T*, long, int, T*);                                   \
    long oracle_hausdorff_##SUF(const T*, long, const T*, long, int, T*);                                       \
    long oracle_distance_stats_##SUF(const T*, long, const T*, long, int, T*);                                  \
    long oracle_generate_##SUF(int, long, int, const double*, double, uint64_t, uint32_t, T*);                  \
    long oracle_order_##SUF(const T*, long, int, T, long*, long*);                                              \
    long oracle_tree_##SUF(int, const T*, long, int, T, long, const long*, long*, long, long*, long, long*, long, \
                           T*);
DECL(f32, float)
DECL(f64, double)
const char* oracle_last_error();
}

static int g_fail = 0;
#define EXPECT(c)                                                             \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #c); \
            g_fail = 1;                                                       \
        }                                                                     \
    } while (0)

// 2-D spiral / 3-D helix from a 64-bit LCG and Box-Muller
static uint64_t g_s = 0x9e3779b97f4a7c15ull;
static double uni() {
    g_s = g_s * 6364136223846793005ull + 1442695040888963407ull;
    return ((g_s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
static double gauss() { return std::sqrt(-2 * std::log(uni())) * std::cos(6.283185307179586 * uni()); }

template <typename T>
static std::vector<T> cloud(long n, int dim) {
    std::vector<T> P((size_t)n * dim);
    for (long i = 0; i < n; ++i) {
        const double t = uni() * 6 * 3.141592653589793;
        const double c[3] = {dim == 2 ? 0.5 * t * std::cos(t) : std::cos(t), dim == 2 ? 0.5 * t * std::sin(t) : std::sin(t),
                             0.15 * t};
        for (int d = 0; d < dim; ++d) P[(size_t)i * dim + d] = (T)((d < 3 ? c[d] : 0.0) + 0.05 * gauss());
    }
    return P;
}

#define RUN(SUF, T)                                                                                               \
    static void run_##SUF() {                                                                                     \
        for (int dim : {2, 3}) {                                                                                  \
            const long n = 3000;                                                                                  \
            const T r1 = dim == 2 ? T(0.45) : T(0.2), r2 = 2 * r1;                                                \
            std::vector<T> P = cloud<T>(n, dim), S((size_t)n * dim), S2((size_t)n * dim);                         \
            std::vector<long> idx(n), kept(n), order(n), off(n + 1);                                               \
            long K = oracle_choose_##SUF(P.data(), n, dim, r1, S.data(), idx.data());                             \
            EXPECT(K > 0 && K <= n);                                                                              \
            std::vector<int> labels(n), counts(K);                                                                \
            std::vector<T> sums((size_t)K * dim);                                                                 \
            std::vector<unsigned char> moved((size_t)K * dim);                                                    \
            S2 = S;                                                                                               \
            EXPECT(oracle_evolve_iter_##SUF(P.data(), n, S2.data(), K, dim, r1, labels.data(), counts.data(),     \
                                            sums.data(), moved.data()) >= 0);                                     \
            S2 = S;                                                                                               \
            EXPECT(oracle_evolve_omp_iter_##SUF(P.data(), n, S2.data(), K, dim, r1, counts.data(), sums.data()) >= 0); \
            S2 = S;                                                                                               \
            EXPECT(oracle_evolve_##SUF(P.data(), n, S2.data(), K, dim, r1, 0) > 0);                               \
            std::vector<T> SD = S2;                                                                               \
            EXPECT(oracle_decimate_##SUF(SD.data(), K, dim, r2, 0, kept.data()) >= 0);                            \
            SD = S2;                                                                                              \
            EXPECT(oracle_decimate_##SUF(SD.data(), K, dim, r2, 1, kept.data()) >= 0);                            \
            long rounds = 0;                                                                                      \
            const long KR = oracle_rd_##SUF(P.data(), n, dim, r1, r2, 0, S2.data(), &rounds);                     \
            EXPECT(KR > 0 && rounds > 0);                                                                         \
            EXPECT(oracle_rd_own_##SUF(P.data(), n, n / 2, dim, r1, r2, 0, SD.data(), &rounds) > 0);              \
            const long nc = oracle_order_##SUF(S2.data(), KR, dim, r2, order.data(), off.data());                 \
            EXPECT(nc > 0 && off[(size_t)nc] == KR);                                                              \
            T h = 0, st[4];                                                                                       \
            std::vector<T> dist((size_t)KR);                                                                      \
            EXPECT(oracle_hausdorff_##SUF(S2.data(), KR, P.data(), n, dim, &h) >= 0 && h >= 0);                   \
            EXPECT(oracle_distance_stats_##SUF(S2.data(), KR, P.data(), n, dim, st) >= 0);                        \
            EXPECT(oracle_set_distances_##SUF(S2.data(), KR, P.data(), n, dim, dist.data()) >= 0);                \
            /* frozen-membership tape, replay and VJP */                                                          \
            const long iters = 5;                                                                                 \
            std::vector<int> tl((size_t)iters * n), tc((size_t)iters * K);                                        \
            std::vector<unsigned char> tm((size_t)iters * K * dim);                                               \
            S2 = S;                                                                                               \
            EXPECT(oracle_evolve_tape_##SUF(P.data(), n, S2.data(), K, dim, r1, iters, tl.data(), tc.data(),      \
                                            tm.data()) >= 0);                                                     \
            std::vector<T> S3 = S;                                                                                \
            EXPECT(oracle_evolve_replay_##SUF(P.data(), n, S3.data(), K, dim, iters, tl.data(), tm.data()) >= 0); \
            EXPECT(S3 == S2);                                                                                     \
            std::vector<T> G((size_t)K * dim, T(1)), gp((size_t)n * dim);                                         \
            EXPECT(oracle_evolve_vjp_##SUF(n, K, dim, iters, tl.data(), tc.data(), tm.data(), G.data(),           \
                                           gp.data()) >= 0);                                                      \
            /* trees and every tiled driver, with and without refinement */                                      \
            const long tpd[3] = {3, 3, 3};                                                                        \
            const long cap = 400;                                                                                 \
            for (int kind = 0; kind < 3; ++kind) {                                                                \
                const long L = oracle_tree_##SUF(kind, P.data(), n, dim, r1, cap, tpd, nullptr, 0, nullptr, 0,    \
                                                 nullptr, 0, nullptr);                                            \
                EXPECT(L > 0);                                                                                    \
                if (L <= 0) continue;                                                                             \
                std::vector<long> info((size_t)L * 4), own(n + 1), nb((size_t)64 * n);                            \
                std::vector<T> bounds((size_t)L * 2 * dim);                                                       \
                EXPECT(oracle_tree_##SUF(kind, P.data(), n, dim, r1, cap, tpd, info.data(), L, own.data(), n + 1, \
                                         nb.data(), (long)nb.size(), bounds.data()) == L);                        \
            }                                                                                                     \
            std::vector<T> ST((size_t)4 * n * dim + 64);                                                          \
            for (int alg = 0; alg <= 6; ++alg)                                                                    \
                for (int ref = 0; ref < 2; ++ref)                                                                 \
                    EXPECT(oracle_tiled_##SUF(alg, P.data(), n, dim, r1, r2, cap, tpd, ref, 0, ST.data(),         \
                                              4 * n + 16) >= 0);                                                  \
        }                                                                                                         \
        std::vector<T> out(500 * 4);                                                                              \
        const double prm[4] = {12.0, 0.5, 0.0, 0.0};                                                              \
        for (int kind = 0; kind < 8; ++kind) {                                                                    \
            const int dim = kind == 0 || kind == 2 ? 2 : kind == 1 || kind == 3 || kind == 5 ? 3 : 4;             \
            const double poly[1 + 2 * 4] = {1.0, 0, 0, 0, 0, 1, 1, 1, 1};                                         \
            EXPECT(oracle_generate_##SUF(kind, 500, dim, kind == 6 ? poly : prm, 0.01, 7, 3, out.data()) >= 0);   \
        }                                                                                                         \
    }

RUN(f32, float)
RUN(f64, double)

int main() {
    run_f32();
    run_f64();
    if (g_fail) std::fprintf(stderr, "last oracle error: %s\n", oracle_last_error());
    std::printf(g_fail ? "oracle self-test FAILED\n" : "oracle self-test ok\n");
    return g_fail;
}
