// rdhip/cpu_ridge_detection.hpp -- header-only drop-in for the reference's CPU entry points,
// run on the GPU through the C ABI of librdhip.so (include/rdhip.h):
//   rd::RidgeDetection<T>
//       (ridge_detection/rd/cpu/brute_force/ridge_detection.hpp:57-117, ridge_detection.inl:30-286)
//   rd::tiled::rdTiledData<T>, rd::tiled::TiledRDAlgorithm, rd::tiled::tiledRidgeDetection<ALG, T>
//       (ridge_detection/rd/cpu/tiled/rd_tiled.hpp:55-110, 1763-1798)
// A program using them includes this header instead and links -lrdhip.  Class members, method
// signatures, enum values and the returned (whole, build, rdTiles, refinement) times are the
// reference's; results equal the reference's serial build (the OpenMP build sums in a
// thread-dependent order).  Differences:
//   * ompSetNumThreads / noOMP and cpuThreads are accepted and have no effect;
//   * verbose and drawTiles print and draw nothing;
//   * failures throw std::runtime_error with the library's message (the reference has no
//     error path);
//   * times are the library's wall-clock milliseconds of the GPU run.
#ifndef RDHIP_CPU_RIDGE_DETECTION_HPP
#define RDHIP_CPU_RIDGE_DETECTION_HPP

#include "../rdhip.h"

#include <cstddef>
#include <deque>
#include <list>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace rd {

namespace rdhip_cpu_detail {
inline void check(int rc, const char* where) {
    if (rc != RDHIP_OK) throw std::runtime_error(std::string(where) + ": " + rdhip_last_error());
}
template <typename T>
constexpr int dtype() {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "T must be float or double");
    return std::is_same<T, float>::value ? RDHIP_F32 : RDHIP_F64;
}
// One context for the duration of a call.
struct Ctx {
    rdhip_ctx* c = nullptr;
    Ctx(int dtype, int dim) { check(rdhip_create(&c, dtype, dim, nullptr), "rdhip_create"); }
    ~Ctx() { rdhip_destroy(c); }
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;
};
}  // namespace rdhip_cpu_detail

// Brute-force ridge detection (choose, then rounds of evolve + decimate until the number of
// chosen points is stable).  S must hold np x dim values (ROW_MAJOR); ns_ receives the count.
// order_ = true also fills chainList_ with the ridge ordered into chains (rd::orderSamples
// with |p - q|^2 <= r2^2), as pointers into S.
template <typename T>
class RidgeDetection {
public:
    size_t ns_;
    std::list<std::deque<T const*>> chainList_;
    bool verbose_;
    bool order_;

    RidgeDetection() : ns_(0), verbose_(false), order_(false), ompNumThreads_(0), noOMP_(false) {}
    RidgeDetection(const RidgeDetection&) = delete;
    RidgeDetection& operator=(const RidgeDetection&) = delete;

    void ompSetNumThreads(int t) { ompNumThreads_ = t; }
    void noOMP() {
        ompNumThreads_ = 0;
        noOMP_ = true;
    }

    // P: np points, ROW_MAJOR
    void ridgeDetection(T const* P, size_t np, T* S, T r1, T r2, size_t dim) {
        run(P, np, np, S, r1, r2, dim);
    }

    // Choose from samples[0] only, evolve over all sample sets in order (a tile's own samples
    // followed by its neighbours').
    void ridgeDetection(std::vector<std::pair<T const*, size_t>> samples, T* S, T r1, T r2, size_t dim) {
        if (samples.empty()) throw std::runtime_error("ridgeDetection: no sample sets");
        if (samples.size() == 1) return run(samples[0].first, samples[0].second, samples[0].second, S, r1, r2, dim);
        size_t n = 0;
        for (auto& s : samples) n += s.second;
        std::vector<T> all;
        all.reserve(n * dim);
        for (auto& s : samples) all.insert(all.end(), s.first, s.first + s.second * dim);
        run(all.data(), n, samples[0].second, S, r1, r2, dim);
    }

protected:
    int ompNumThreads_;
    bool noOMP_;

private:
    void run(T const* P, size_t n, size_t nOwn, T* S, T r1, T r2, size_t dim) {
        rdhip_cpu_detail::Ctx x(rdhip_cpu_detail::dtype<T>(), (int)dim);
        rdhip_cpu_detail::check(rdhip_set_points(x.c, P, (int64_t)n, RDHIP_ROW_MAJOR, 0, 0), "rdhip_set_points");
        int64_t K = 0;
        rdhip_cpu_detail::check(rdhip_ridge_detection_own(x.c, (int64_t)nOwn, (double)r1, (double)r2, 0, &K),
                                "rdhip_ridge_detection_own");
        ns_ = (size_t)K;
        if (K > 0) rdhip_cpu_detail::check(rdhip_get_points(x.c, S, RDHIP_ROW_MAJOR, 0, 0), "rdhip_get_points");
        chainList_.clear();
        if (!order_ || K == 0) return;
        std::vector<int32_t> order((size_t)K);
        std::vector<int64_t> offsets((size_t)K + 1);
        int64_t chains = 0;
        rdhip_cpu_detail::check(rdhip_order_chains(x.c, (double)r2, order.data(), offsets.data(), &chains),
                                "rdhip_order_chains");
        for (int64_t c = 0; c < chains; ++c) {
            std::deque<T const*> chain;
            for (int64_t i = offsets[(size_t)c]; i < offsets[(size_t)c + 1]; ++i)
                chain.push_back(S + (size_t)order[(size_t)i] * dim);
            chainList_.push_back(std::move(chain));
        }
    }
};

namespace tiled {

// rd/cpu/tiled/rd_tiled.hpp:55-100
template <typename T>
struct rdTiledData {
    size_t dim;
    size_t np;
    size_t ns;
    size_t maxTileCapacity;
    std::vector<size_t> nTilesPerDim;
    T r1;
    T r2;
    T a, b, s;
    T const* P;
    std::vector<T> S;

    rdTiledData() : dim(0), np(0), ns(0), maxTileCapacity(0), nTilesPerDim(), r1(0), r2(0), a(0), b(0), s(0), P(nullptr), S() {}
};

// rd/cpu/tiled/rd_tiled.hpp:102-110 (values equal enum rdhip_tiled_algorithm's)
enum TiledRDAlgorithm {
    TILED_LOCAL_TREE_MIXED_RD,
    TILED_GROUPED_TREE_MIXED_RD,
    TILED_LOCAL_TREE_LOCAL_RD,
    TILED_LOCAL_LIST,
    TILED_GROUPED_TREE_LOCAL_RD,
    EXT_TILE_TREE_MIXED_RD
};

// Tiled ridge detection of dataPack.P (np x dim, ROW_MAJOR) with dataPack's radii, tile
// capacity and initial tiles per dimension; fills dataPack.S (ns x dim) and dataPack.ns and
// returns (whole, build tree, ridge detection of the tiles, end-phase refinement) in ms.
// TILED_LOCAL_LIST is a stub in the reference; here as there it returns zeros and no points.
template <TiledRDAlgorithm ALGORITHM, typename T>
std::tuple<float, float, float, float> tiledRidgeDetection(rdTiledData<T>& dataPack, int cpuThreads = 1,
                                                          bool endPhaseRefinement = false, bool verbose = false,
                                                          bool drawTiles = false) {
    (void)cpuThreads;
    (void)verbose;
    (void)drawTiles;
    if (dataPack.nTilesPerDim.size() != dataPack.dim)
        throw std::runtime_error("tiledRidgeDetection: nTilesPerDim must hold dim values");
    dataPack.ns = 0;
    if (ALGORITHM == TILED_LOCAL_LIST) return std::make_tuple(0.f, 0.f, 0.f, 0.f);
    rdhip_cpu_detail::Ctx x(rdhip_cpu_detail::dtype<T>(), (int)dataPack.dim);
    rdhip_cpu_detail::check(rdhip_set_points(x.c, dataPack.P, (int64_t)dataPack.np, RDHIP_ROW_MAJOR, 0, 0),
                            "rdhip_set_points");
    std::vector<int64_t> tpd(dataPack.nTilesPerDim.begin(), dataPack.nTilesPerDim.end());
    int64_t K = 0;
    rdhip_cpu_detail::check(rdhip_tiled_ridge_detection(x.c, (int)ALGORITHM, (double)dataPack.r1, (double)dataPack.r2,
                                                        (int64_t)dataPack.maxTileCapacity, tpd.data(),
                                                        endPhaseRefinement ? 1 : 0, 0, &K),
                            "rdhip_tiled_ridge_detection");
    dataPack.ns = (size_t)K;
    dataPack.S.assign((size_t)K * dataPack.dim, T(0));
    if (K > 0)
        rdhip_cpu_detail::check(rdhip_get_points(x.c, dataPack.S.data(), RDHIP_ROW_MAJOR, 0, 0), "rdhip_get_points");
    rdhip_timers t{};
    rdhip_cpu_detail::check(rdhip_get_timers(x.c, &t), "rdhip_get_timers");
    const double tiles = t.whole_ms - t.build_ms - t.refine_ms;
    return std::make_tuple((float)t.whole_ms, (float)t.build_ms, (float)(tiles > 0 ? tiles : 0), (float)t.refine_ms);
}

}  // namespace tiled
}  // namespace rd

#endif  // RDHIP_CPU_RIDGE_DETECTION_HPP
