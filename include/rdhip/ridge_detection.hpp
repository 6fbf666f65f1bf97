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
//   * the device facades DeviceChoose / DeviceEvolve / DeviceDecimate (namespace
//     rd::gpu::bruteForce, ridge_detection/rd/gpu/device/device_{choose,evolve,decimate}.cuh)
//     return hipError_t; the library's message for a failure is rdhip_last_error().  They run
//     synchronously on the given stream, each call on a context of its own;
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
#include <vector>

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
    int* dNs_;         // device copy of ns_ (the reference's rdBruteForceNs symbol)
    // The reference's evolve / decimate work buffers.  The library keeps its own, so these
    // stay null; they are declared for source compatibility.
    T* dCordSums_ = nullptr;
    size_t csPitch_ = 0;
    int* dSpherePointCnt_ = nullptr;
    T* dDistMtx_ = nullptr;
    size_t distMtxPitch_ = 0;
    char* dPtsMask_ = nullptr;
    bool verbose_;
    long long maxEvolveIters_ = 0;   // 0: evolve to convergence (reference behaviour)

    RidgeDetection(size_t np, T r1, T r2, bool verbose = false)
        : np_(np), ns_(0), r1_(r1), r2_(r2), dP_(nullptr), pPitch_(0), dS_(nullptr), sPitch_(0), dNs_(nullptr),
          verbose_(verbose) {
        static_assert(IN_MEM_LAYOUT != SOA && OUT_MEM_LAYOUT != SOA, "ROW_MAJOR or COL_MAJOR (as in the reference)");
        try {
            detail::check(rdhip_create(&ctx_, detail::dtype<T>(), DIM, nullptr), "rdhip_create");
            detail::checkHip(hipMalloc((void**)&dNs_, sizeof(int)), "hipMalloc");
            detail::checkHip(hipMemset(dNs_, 0, sizeof(int)), "hipMemset");
            if (IN_MEM_LAYOUT == COL_MAJOR)
                detail::checkHip(hipMallocPitch((void**)&dP_, &pPitch_, np_ * sizeof(T), DIM), "hipMallocPitch");
            else
                detail::checkHip(hipMalloc((void**)&dP_, np_ * DIM * sizeof(T)), "hipMalloc");
            if (OUT_MEM_LAYOUT == COL_MAJOR)
                detail::checkHip(hipMallocPitch((void**)&dS_, &sPitch_, np_ * sizeof(T), DIM), "hipMallocPitch");
            else
                detail::checkHip(hipMalloc((void**)&dS_, np_ * DIM * sizeof(T)), "hipMalloc");
        } catch (...) {
            release();
            throw;
        }
    }
    virtual ~RidgeDetection() { release(); }
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
        const int ns32 = (int)ns_;
        detail::checkHip(hipMemcpy(dNs_, &ns32, sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
    }
    // ns_ from the device copy dNs_ (simulation.inl: getChosenSamplesCount)
    void getChosenSamplesCount() {
        int ns = 0;
        detail::checkHip(hipMemcpy(&ns, dNs_, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
        ns_ = (size_t)ns;
    }
    rdhip_ctx* context() const { return ctx_; }

private:
    rdhip_ctx* ctx_ = nullptr;

    void release() {
        if (dP_) (void)hipFree(dP_);
        if (dS_) (void)hipFree(dS_);
        if (dNs_) (void)hipFree(dNs_);
        dP_ = dS_ = nullptr;
        dNs_ = nullptr;
        rdhip_destroy(ctx_);
        ctx_ = nullptr;
    }
};

// ---- device-level facades ------------------------------------------------------------
// The reference's DeviceChoose / DeviceEvolve / DeviceDecimate structs with their template
// parameters and argument lists.  Pointers are device memory; strides count elements between
// consecutive points (ROW_MAJOR, usually DIM) or between coordinate rows (COL_MAJOR, the pitch
// in elements).  Results follow the reference's serial CPU path (choose order, first-index
// ties, summation in sample order), so they do not vary between runs.
namespace facade_detail {
inline hipError_t toHip(int rc) {
    return rc == RDHIP_OK ? hipSuccess : rc == RDHIP_ERR_INVALID ? hipErrorInvalidValue : hipErrorUnknown;
}
// A context on `stream` for one facade call (destroyed, after a stream synchronisation, on exit).
struct ScopedCtx {
    rdhip_ctx* c = nullptr;
    int rc = RDHIP_OK;
    ScopedCtx(int dtype, int dim, hipStream_t stream) { rc = rdhip_create(&c, dtype, dim, (void*)stream); }
    ~ScopedCtx() { rdhip_destroy(c); }
    ScopedCtx(const ScopedCtx&) = delete;
    ScopedCtx& operator=(const ScopedCtx&) = delete;
};
}  // namespace facade_detail

// rd/gpu/device/device_choose.cuh:55-82: greedy choose of r-separated points from
// d_inputPoints into d_chosenPoints (OUT_MEM_LAYOUT); the count goes to *d_chosenPointsNum.
struct DeviceChoose {
    template <int DIM, DataMemoryLayout IN_MEM_LAYOUT, DataMemoryLayout OUT_MEM_LAYOUT, typename T>
    static hipError_t choose(T const* d_inputPoints, T* d_chosenPoints, int inputPointsNum, int* d_chosenPointsNum,
                             T r, int inPtsStride, int chosenPtsStride, hipStream_t stream = nullptr,
                             bool debugSynchronous = false) {
        (void)debugSynchronous;
        facade_detail::ScopedCtx x(gpu::detail::dtype<T>(), DIM, stream);
        int rc = x.rc;
        int64_t K = 0;
        if (rc == RDHIP_OK)
            rc = rdhip_set_points(x.c, d_inputPoints, inputPointsNum, gpu::detail::layout(IN_MEM_LAYOUT), inPtsStride, 1);
        if (rc == RDHIP_OK) rc = rdhip_choose(x.c, (double)r, &K, nullptr);
        if (rc == RDHIP_OK && K > 0)
            rc = rdhip_get_points(x.c, d_chosenPoints, gpu::detail::layout(OUT_MEM_LAYOUT), chosenPtsStride, 1);
        if (rc != RDHIP_OK) return facade_detail::toHip(rc);
        const int k32 = (int)K;
        hipError_t e = hipMemcpyAsync(d_chosenPointsNum, &k32, sizeof(int), hipMemcpyHostToDevice, stream);
        return e != hipSuccess ? e : hipStreamSynchronize(stream);
    }
};

// rd/gpu/device/device_evolve.cuh:29-97: evolve d_chosenSamples (in place, OUTPUT_MEM_LAYOUT)
// until no coordinate moves.  d_cordSums (OUTPUT_MEM_LAYOUT, stride cordSumsStride) and
// d_spherePointCount receive the last iteration's coordinate sums and member counts, as the
// reference's work buffers hold them on return.  The second overload evolves over the samples
// followed by the neighbour samples (a tile's own and neighbouring points).
struct DeviceEvolve {
    template <int DIM, DataMemoryLayout INPUT_MEM_LAYOUT, DataMemoryLayout OUTPUT_MEM_LAYOUT, typename T>
    static hipError_t evolve(T const* d_inputPoints, T* d_chosenSamples, T* d_cordSums, int* d_spherePointCount,
                             int inputPointsNum, int chosenPointsNum, T r, int inPtsStride, int chosenPtsStride,
                             int cordSumsStride, hipStream_t stream = nullptr, bool debugSynchronous = false) {
        (void)debugSynchronous;
        facade_detail::ScopedCtx x(gpu::detail::dtype<T>(), DIM, stream);
        int rc = x.rc;
        if (rc == RDHIP_OK)
            rc = rdhip_set_points(x.c, d_inputPoints, inputPointsNum, gpu::detail::layout(INPUT_MEM_LAYOUT),
                                  inPtsStride, 1);
        return finish<DIM, OUTPUT_MEM_LAYOUT>(x, rc, d_chosenSamples, d_cordSums, d_spherePointCount, chosenPointsNum,
                                              r, chosenPtsStride, cordSumsStride, stream);
    }

    template <int DIM, DataMemoryLayout INPUT_MEM_LAYOUT, DataMemoryLayout OUTPUT_MEM_LAYOUT, typename T>
    static hipError_t evolve(T const* d_inputPoints, T const* d_neighbourPoints, T* d_chosenSamples, T* d_cordSums,
                             int* d_spherePointCount, int inputPointsNum, int neighbourPointsNum, int chosenPointsNum,
                             T r, int inPtsStride, int neighbourPtsStride, int chosenPtsStride, int cordSumsStride,
                             hipStream_t stream = nullptr, bool debugSynchronous = false) {
        (void)debugSynchronous;
        // own samples, then the neighbour samples, in one buffer of the input layout
        const size_t n = (size_t)inputPointsNum + (size_t)neighbourPointsNum;
        T* all = nullptr;
        hipError_t e = hipMalloc((void**)&all, sizeof(T) * (n > 0 ? n : 1) * DIM);
        if (e != hipSuccess) return e;
        const bool row = INPUT_MEM_LAYOUT == ROW_MAJOR;
        const size_t rowB = sizeof(T) * DIM;
        if (inputPointsNum > 0)
            e = row ? hipMemcpy2DAsync(all, rowB, d_inputPoints, sizeof(T) * inPtsStride, rowB, inputPointsNum,
                                       hipMemcpyDeviceToDevice, stream)
                    : hipMemcpy2DAsync(all, sizeof(T) * n, d_inputPoints, sizeof(T) * inPtsStride,
                                       sizeof(T) * inputPointsNum, DIM, hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess && neighbourPointsNum > 0)
            e = row ? hipMemcpy2DAsync(all + (size_t)inputPointsNum * DIM, rowB, d_neighbourPoints,
                                       sizeof(T) * neighbourPtsStride, rowB, neighbourPointsNum, hipMemcpyDeviceToDevice,
                                       stream)
                    : hipMemcpy2DAsync(all + inputPointsNum, sizeof(T) * n, d_neighbourPoints,
                                       sizeof(T) * neighbourPtsStride, sizeof(T) * neighbourPointsNum, DIM,
                                       hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
            (void)hipFree(all);
            return e;
        }
        facade_detail::ScopedCtx x(gpu::detail::dtype<T>(), DIM, stream);
        int rc = x.rc;
        if (rc == RDHIP_OK)
            rc = rdhip_set_points(x.c, all, (int64_t)n, row ? RDHIP_ROW_MAJOR : RDHIP_SOA, 0, 1);
        (void)hipFree(all);   // set_points copied the samples
        return finish<DIM, OUTPUT_MEM_LAYOUT>(x, rc, d_chosenSamples, d_cordSums, d_spherePointCount, chosenPointsNum,
                                              r, chosenPtsStride, cordSumsStride, stream);
    }

private:
    template <int DIM, DataMemoryLayout OUT, typename T>
    static hipError_t finish(facade_detail::ScopedCtx& x, int rc, T* dS, T* dSums, int* dCounts, int ns, T r,
                             int sStride, int csStride, hipStream_t stream) {
        const int lay = gpu::detail::layout(OUT);
        if (rc == RDHIP_OK) rc = rdhip_set_chosen_ex(x.c, dS, ns, lay, sStride, 1);
        if (rc == RDHIP_OK) rc = rdhip_evolve(x.c, (double)r, 0, nullptr);
        if (rc == RDHIP_OK && ns > 0) rc = rdhip_get_points(x.c, dS, lay, sStride, 1);
        if (rc != RDHIP_OK || ns == 0) return facade_detail::toHip(rc);
        // the converged set is a fixed point: one more pass yields the last pass's sums and counts
        std::vector<int> counts((size_t)ns);
        std::vector<T> sums((size_t)ns * DIM);
        int flag = 0;
        rc = rdhip_evolve_iteration(x.c, (double)r, nullptr, counts.data(), sums.data(), nullptr, &flag);
        if (rc != RDHIP_OK) return facade_detail::toHip(rc);
        hipError_t e = hipMemcpyAsync(dCounts, counts.data(), sizeof(int) * ns, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess && dSums) {
            if (OUT == ROW_MAJOR)
                e = hipMemcpy2DAsync(dSums, sizeof(T) * csStride, sums.data(), sizeof(T) * DIM, sizeof(T) * DIM, ns,
                                     hipMemcpyHostToDevice, stream);
            else {
                std::vector<T> cm((size_t)ns * DIM);
                for (int k = 0; k < ns; ++k)
                    for (int d = 0; d < DIM; ++d) cm[(size_t)d * ns + k] = sums[(size_t)k * DIM + d];
                e = hipMemcpy2DAsync(dSums, sizeof(T) * csStride, cm.data(), sizeof(T) * ns, sizeof(T) * ns, DIM,
                                     hipMemcpyHostToDevice, stream);
            }
        }
        return e != hipSuccess ? e : hipStreamSynchronize(stream);
    }
};

// rd/gpu/device/device_decimate.cuh:50-117: decimation of d_chosenPoints (in place, stable)
// with radius r; *chosenPointsNum (device memory) holds the count on entry and on return.
// decimateDistMtx applies the same rule; its distance-matrix and mask work buffers are not
// needed and are left untouched.
struct DeviceDecimate {
    template <int DIM, DataMemoryLayout MEM_LAYOUT, typename T>
    static hipError_t decimate(T* d_chosenPoints, int* chosenPointsNum, T r, int stride, hipStream_t stream = nullptr,
                               bool debug_synchronous = false) {
        (void)debug_synchronous;
        int K = 0;
        hipError_t e = hipMemcpyAsync(&K, chosenPointsNum, sizeof(int), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return e;
        if (K <= 0) return hipSuccess;
        facade_detail::ScopedCtx x(gpu::detail::dtype<T>(), DIM, stream);
        const int lay = gpu::detail::layout(MEM_LAYOUT);
        int rc = x.rc;
        int64_t left = 0;
        if (rc == RDHIP_OK) rc = rdhip_set_points(x.c, d_chosenPoints, K, lay, stride, 1);
        if (rc == RDHIP_OK) rc = rdhip_set_chosen_ex(x.c, d_chosenPoints, K, lay, stride, 1);
        if (rc == RDHIP_OK) rc = rdhip_decimate(x.c, (double)r, RDHIP_DECIMATE, nullptr, &left);
        if (rc == RDHIP_OK && left > 0) rc = rdhip_get_points(x.c, d_chosenPoints, lay, stride, 1);
        if (rc != RDHIP_OK) return facade_detail::toHip(rc);
        const int k32 = (int)left;
        e = hipMemcpyAsync(chosenPointsNum, &k32, sizeof(int), hipMemcpyHostToDevice, stream);
        return e != hipSuccess ? e : hipStreamSynchronize(stream);
    }

    template <int DIM, DataMemoryLayout MEM_LAYOUT, typename T>
    static hipError_t decimateDistMtx(T* d_chosenPoints, int* chosenPointsNum, int chPtsStride, T* d_distMtx,
                                      int distMtxStride, char* d_pointsMask, T r, hipStream_t stream = nullptr,
                                      bool debug_synchronous = false) {
        (void)d_distMtx;
        (void)distMtxStride;
        (void)d_pointsMask;
        return decimate<DIM, MEM_LAYOUT, T>(d_chosenPoints, chosenPointsNum, r, chPtsStride, stream, debug_synchronous);
    }
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
    // inputPoints: host memory, ROW_MAJOR (as in the reference, whatever IN_MEM_LAYOUT, which
    // only names the device copy's layout); SOA throws std::runtime_error as the reference does
    // (rd/gpu/device/tiled/simulation.cuh:100-117)
    RidgeDetection(int pointsNum, T const* inputPoints, bool enableTiming = false, bool debugSynchronous = false)
        : pointsNum_(pointsNum), enableTiming_(enableTiming), debugSynchronous_(debugSynchronous) {
        if (IN_MEM_LAYOUT != ROW_MAJOR && IN_MEM_LAYOUT != COL_MAJOR) throw std::runtime_error("Unsupported memory layout!");
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
    // chosenPoints: host memory for getChosenPointsNum() points, always written ROW_MAJOR
    // (OUT_MEM_LAYOUT names the device layout, which the reference transposes on the way
    // back); SOA throws std::runtime_error (rd/gpu/device/tiled/simulation.cuh:174-203)
    void getChosenPoints(T* chosenPoints) {
        if (OUT_MEM_LAYOUT != ROW_MAJOR && OUT_MEM_LAYOUT != COL_MAJOR) throw std::runtime_error("Unsupported memory layout!");
        if (chosenPointsNum_ == 0) return;
        gpu::detail::check(rdhip_get_points(ctx_, chosenPoints, RDHIP_ROW_MAJOR, 0, 0), "rdhip_get_points");
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
