// rdhip/ridge_detection.hpp -- header-only drop-in for the reference's GPU classes, over the
// C ABI of librdhip.so (include/rdhip.h).  A program using the reference's
//   rd::gpu::bruteForce::RidgeDetection<T, DIM, IN, OUT>
//       (ridge_detection/rd/gpu/device/brute_force/simulation.cuh:35-112)
//   rd::gpu::tiled::RidgeDetection<DIM, IN, OUT, ALG, POLICY, TILE, T>
//       (ridge_detection/rd/gpu/device/tiled/simulation.cuh:33-241)
// includes this header instead and links -lrdhip.  Template parameters, constructor and
// method signatures, public members and enum values are the reference's.  Differences:
//   * errors throw rd::gpu::Error (the reference calls checkCudaErrors, which exits);
//   * TileType gains RD_LOCAL_TILE and RD_GROUPED_TILE (the reference GPU only has
//     RD_EXTENDED_TILE; the CPU reference has all three trees);
//   * RD_KD_TREE / RD_BVH_TREE are declared but, as in the reference, not implemented;
//   * rd::gpu::SamplesGenerator<T> (ridge_detection/rd/gpu/device/samples_generator.cuh:
//     291-384) draws from librdhip's counter-based generator (seed 1234, as the reference's
//     curand_init), so its samples follow the reference's distributions but are not the
//     reference's curand values; the T** (one pointer per coordinate) overloads are not
//     provided.
#ifndef RDHIP_RIDGE_DETECTION_HPP
#define RDHIP_RIDGE_DETECTION_HPP

#include "../rdhip.h"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstddef>
#include <initializer_list>
#include <stdexcept>
#include <string>
#include <type_traits>

namespace rd {

// rd/utils/memory.h:22-27
enum DataMemoryLayout { ROW_MAJOR, COL_MAJOR, SOA };

namespace gpu {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

namespace detail {
inline void check(int rc, const char* where) {
    if (rc != RDHIP_OK) throw Error(rc, std::string(where) + ": " + rdhip_last_error());
}
inline void checkHip(hipError_t e, const char* where) {
    if (e != hipSuccess) throw Error(RDHIP_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}
template <typename T>
constexpr int dtype() {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "T must be float or double");
    return std::is_same<T, float>::value ? RDHIP_F32 : RDHIP_F64;
}
inline int layout(DataMemoryLayout l) { return l == ROW_MAJOR ? RDHIP_ROW_MAJOR : l == COL_MAJOR ? RDHIP_COL_MAJOR : RDHIP_SOA; }

// rd/gpu/agent/agent_tiled_ridge_detection.cuh:96-102 (times in milliseconds here)
struct TiledRidgeDetectionTimers {
    long long int rdTilesTime;
    long long int refinementTime;
    long long int wholeTime;
};
}  // namespace detail

// rd/gpu/device/dispatch/dispatch_ridge_detection.cuh:39-74
enum RidgeDetectionAlgorithm { RD_BRUTE_FORCE, RD_KD_TREE, RD_BVH_TREE };

// rd/gpu/device/samples_generator.cuh:291-384: noisy samples of simple shapes written into a
// device buffer (ROW_MAJOR: n x dim; COL_MAJOR: dim rows of stride n).  As in the reference,
// sigma is divided by sqrt(dim) so that a sample's distance to the shape has spread sigma.
template <typename T>
struct SamplesGenerator {
    template <DataMemoryLayout SAMPLES_ORDER = ROW_MAJOR>
    static void spiral2D(size_t n, T a, T b, T sigma, T* dSamples) {
        gen(SAMPLES_ORDER, RDHIP_GEN_SPIRAL2D, n, 2, {(double)a, (double)b}, sigma, dSamples);
    }
    template <DataMemoryLayout SAMPLES_ORDER = ROW_MAJOR>
    static void spiral3D(size_t n, T a, T b, T sigma, T* dSamples) {
        gen(SAMPLES_ORDER, RDHIP_GEN_SPIRAL3D, n, 3, {(double)a, (double)b}, sigma, dSamples);
    }
    template <DataMemoryLayout SAMPLES_ORDER = ROW_MAJOR>
    static void segmentND(size_t n, size_t dim, T sigma, T length, T* dSamples, int stride = 0) {
        gen(SAMPLES_ORDER, RDHIP_GEN_SEGMENT, n, (int)dim, {(double)length}, sigma, dSamples,
            SAMPLES_ORDER == ROW_MAJOR ? 0 : stride);
    }
    template <DataMemoryLayout SAMPLES_ORDER = ROW_MAJOR>
    static void circle(size_t n, T x0, T y0, T r, T sigma, T* dSamples) {
        gen(SAMPLES_ORDER, RDHIP_GEN_CIRCLE, n, 2, {(double)x0, (double)y0, (double)r}, sigma, dSamples);
    }
    template <DataMemoryLayout SAMPLES_ORDER = ROW_MAJOR>
    static void sphere(size_t n, T x0, T y0, T z0, T r, T sigma, T* dSamples) {
        gen(SAMPLES_ORDER, RDHIP_GEN_SPHERE, n, 3, {(double)x0, (double)y0, (double)z0, (double)r}, sigma,
            dSamples);
    }

private:
    static void gen(DataMemoryLayout order, int kind, size_t n, int dim, std::initializer_list<double> params,
                    T sigma, T* dSamples, int64_t ld = 0) {
        const double s = (double)sigma / std::sqrt((double)dim);
        detail::check(rdhip_generate(detail::dtype<T>(), kind, (int64_t)n, dim, params.begin(), (int64_t)params.size(),
                                     s, 1234, 0, dSamples, detail::layout(order), ld, 1),
                      "SamplesGenerator");
    }
};

namespace bruteForce {

template <typename T, int DIM, DataMemoryLayout IN_MEM_LAYOUT, DataMemoryLayout OUT_MEM_LAYOUT>
class RidgeDetection {
public:
    size_t np_;        // cardinality of the samples set
    size_t ns_;        // number of chosen points (after ridgeDetection())
    T r1_;             // choose / evolve radius
    T r2_;             // decimate radius
    T* dP_;            // samples in device memory, IN_MEM_LAYOUT; callers write here
    size_t pPitch_;    // COL_MAJOR: row pitch of dP_ in bytes; otherwise 0
    T* dS_;            // chosen points in device memory, OUT_MEM_LAYOUT
    size_t sPitch_;    // COL_MAJOR: row pitch of dS_ in bytes; otherwise 0
    bool verbose_;
    long long maxEvolveIters_ = 0;   // 0: evolve to convergence (reference behaviour)

    RidgeDetection(size_t np, T r1, T r2, bool verbose = false)
        : np_(np), ns_(0), r1_(r1), r2_(r2), dP_(nullptr), pPitch_(0), dS_(nullptr), sPitch_(0), verbose_(verbose) {
        detail::check(rdhip_create(&ctx_, detail::dtype<T>(), DIM, nullptr), "rdhip_create");
        if (IN_MEM_LAYOUT == COL_MAJOR)
            detail::checkHip(hipMallocPitch((void**)&dP_, &pPitch_, np_ * sizeof(T), DIM), "hipMallocPitch");
        else
            detail::checkHip(hipMalloc((void**)&dP_, np_ * DIM * sizeof(T)), "hipMalloc");
        if (OUT_MEM_LAYOUT == COL_MAJOR)
            detail::checkHip(hipMallocPitch((void**)&dS_, &sPitch_, np_ * sizeof(T), DIM), "hipMallocPitch");
        else
            detail::checkHip(hipMalloc((void**)&dS_, np_ * DIM * sizeof(T)), "hipMalloc");
    }
    virtual ~RidgeDetection() {
        if (dP_) (void)hipFree(dP_);
        if (dS_) (void)hipFree(dS_);
        rdhip_destroy(ctx_);
    }
    RidgeDetection(const RidgeDetection&) = delete;
    RidgeDetection& operator=(const RidgeDetection&) = delete;

    // choose, then rounds of evolve + decimate; results in ns_ and dS_
    void ridgeDetection() {
        detail::checkHip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        const int64_t inStride = IN_MEM_LAYOUT == COL_MAJOR ? (int64_t)(pPitch_ / sizeof(T)) : 0;
        detail::check(rdhip_set_points(ctx_, dP_, (int64_t)np_, detail::layout(IN_MEM_LAYOUT), inStride, 1),
                      "rdhip_set_points");
        int64_t ns = 0;
        detail::check(rdhip_ridge_detection(ctx_, (double)r1_, (double)r2_, maxEvolveIters_, &ns),
                      "rdhip_ridge_detection");
        ns_ = (size_t)ns;
        const int64_t outStride = OUT_MEM_LAYOUT == COL_MAJOR ? (int64_t)(sPitch_ / sizeof(T)) : 0;
        if (ns_ > 0)
            detail::check(rdhip_get_points(ctx_, dS_, detail::layout(OUT_MEM_LAYOUT), outStride, 1),
                          "rdhip_get_points");
    }
    void getChosenSamplesCount() {
        int64_t ns = 0;
        detail::check(rdhip_get_count(ctx_, &ns), "rdhip_get_count");
        ns_ = (size_t)ns;
    }
    rdhip_ctx* context() const { return ctx_; }

private:
    rdhip_ctx* ctx_ = nullptr;
};

}  // namespace bruteForce

namespace tiled {

// rd/gpu/agent/agent_tiled_ridge_detection.cuh:44-88, extended with the CPU trees
enum TileType { RD_EXTENDED_TILE, RD_LOCAL_TILE, RD_GROUPED_TILE };
enum TiledRidgeDetectionPolicy { RD_LOCAL, RD_MIXED };

namespace detail {
constexpr int algorithmId(TiledRidgeDetectionPolicy p, TileType t) {
    return t == RD_LOCAL_TILE ? (p == RD_MIXED ? RDHIP_TILED_LOCAL_TREE_MIXED_RD : RDHIP_TILED_LOCAL_TREE_LOCAL_RD)
         : t == RD_GROUPED_TILE ? (p == RD_MIXED ? RDHIP_TILED_GROUPED_TREE_MIXED_RD : RDHIP_TILED_GROUPED_TREE_LOCAL_RD)
         : (p == RD_MIXED ? RDHIP_EXT_TILE_TREE_MIXED_RD : RDHIP_EXT_TILE_TREE_LOCAL_RD);
}
}  // namespace detail

template <int DIM, DataMemoryLayout IN_MEM_LAYOUT, DataMemoryLayout OUT_MEM_LAYOUT,
          RidgeDetectionAlgorithm RD_TILE_ALGORITHM, TiledRidgeDetectionPolicy RD_TILE_POLICY,
          TileType RD_TILE_TYPE, typename T>
class RidgeDetection {
    static_assert(RD_TILE_ALGORITHM == RD_BRUTE_FORCE, "only RD_BRUTE_FORCE is implemented (as in the reference)");

public:
    // inputPoints: host memory, ROW_MAJOR (as in the reference)
    RidgeDetection(int pointsNum, T const* inputPoints, bool enableTiming = false, bool debugSynchronous = false)
        : pointsNum_(pointsNum), enableTiming_(enableTiming), debugSynchronous_(debugSynchronous) {
        gpu::detail::check(rdhip_create(&ctx_, gpu::detail::dtype<T>(), DIM, nullptr), "rdhip_create");
        gpu::detail::check(rdhip_set_points(ctx_, inputPoints, pointsNum, RDHIP_ROW_MAJOR, 0, 0), "rdhip_set_points");
    }
    ~RidgeDetection() { rdhip_destroy(ctx_); }
    RidgeDetection(const RidgeDetection&) = delete;
    RidgeDetection& operator=(const RidgeDetection&) = delete;

    void operator()(T r1, T r2, int maxTileCapacity, int tilesPerDim[DIM], bool endPhaseRefinement) {
        approximate(r1, r2, maxTileCapacity, tilesPerDim, endPhaseRefinement);
    }
    void ridgeDetection(T r1, T r2, int maxTileCapacity, int tilesPerDim[DIM], bool endPhaseRefinement) {
        approximate(r1, r2, maxTileCapacity, tilesPerDim, endPhaseRefinement);
    }
    int getChosenPointsNum() {
        int64_t n = 0;
        gpu::detail::check(rdhip_get_count(ctx_, &n), "rdhip_get_count");
        chosenPointsNum_ = (int)n;
        return chosenPointsNum_;
    }
    // chosenPoints: host memory for getChosenPointsNum() points, in OUT_MEM_LAYOUT
    void getChosenPoints(T* chosenPoints) {
        const int layout = OUT_MEM_LAYOUT == ROW_MAJOR ? RDHIP_ROW_MAJOR : RDHIP_SOA;
        gpu::detail::check(rdhip_get_points(ctx_, chosenPoints, layout, 0, 0), "rdhip_get_points");
    }
    gpu::detail::TiledRidgeDetectionTimers getTimers() {
        rdhip_timers t{};
        gpu::detail::check(rdhip_get_timers(ctx_, &t), "rdhip_get_timers");
        gpu::detail::TiledRidgeDetectionTimers r;
        r.rdTilesTime = (long long)(t.whole_ms - t.refine_ms);
        r.refinementTime = (long long)t.refine_ms;
        r.wholeTime = (long long)t.whole_ms;
        return r;
    }
    rdhip_ctx* context() const { return ctx_; }

private:
    int pointsNum_;
    int chosenPointsNum_ = 0;
    bool enableTiming_;
    bool debugSynchronous_;
    rdhip_ctx* ctx_ = nullptr;

    void approximate(T r1, T r2, int maxTileCapacity, int tilesPerDim[DIM], bool endPhaseRefinement) {
        int64_t tpd[DIM];
        for (int d = 0; d < DIM; ++d) tpd[d] = tilesPerDim[d];
        int64_t n = 0;
        gpu::detail::check(rdhip_tiled_ridge_detection(ctx_, detail::algorithmId(RD_TILE_POLICY, RD_TILE_TYPE),
                                                       (double)r1, (double)r2, maxTileCapacity, tpd,
                                                       endPhaseRefinement ? 1 : 0, 0, &n),
                           "rdhip_tiled_ridge_detection");
        chosenPointsNum_ = (int)n;
    }
};

}  // namespace tiled
}  // namespace gpu
}  // namespace rd

#endif  // RDHIP_RIDGE_DETECTION_HPP
