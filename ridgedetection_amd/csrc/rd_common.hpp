// rd_common.hpp -- shared device/host definitions for the gfx950 ridge-detection kernels.
//
// Layout in HBM (all kernels):
//   * cloud points are SoA: coordinate d of point i lives at P[d * pStride + i];
//     pStride is a multiple of 64 so every dimension row starts 256-byte aligned;
//   * chosen points are SoA too: S[d * sStride + slot].  Every segment (a whole cloud,
//     a tile, or one cloud of a batch) owns the slot range [sOff, sOff + sCap), sOff a
//     multiple of kChosenTile, and slots past its live count K hold a far-away sentinel
//     so that the evolve scan may run over whole tiles of kChosenTile chosen points.
//
// Floating point: every kernel is compiled with -ffp-contract=off.  Distances that decide
// anything bit-exactly are computed as ((0 + t0*t0) + t1*t1) + ... with separate mul/add,
// the order of rd/cpu/brute_force/rd_inner.hpp:41-49; explicit fma() is used only in the
// screening pass of the evolve scan, whose decisions are re-checked exactly.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>

namespace rdhip {

// Fault injection for tests (RDHIP_OPT_INJECT_ALLOC_FAILURE): when positive, the device
// allocation that brings this count to zero fails with hipErrorOutOfMemory.  Per thread.
inline thread_local long long t_allocFailCountdown = 0;

// Every device allocation of the library goes through here.
inline hipError_t deviceMalloc(void** p, size_t bytes) {
    if (t_allocFailCountdown > 0 && --t_allocFailCountdown == 0) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(p, bytes);
}

// The message rdhip_last_error() returns on the calling thread (defined in rd_api.hip).
void set_last_error(const char* msg);

constexpr int kWave = 64;
constexpr int kChosenTile = 16;   // chosen points per screening tile of the evolve scan
constexpr int kMaxDim = 12;

template <typename T> struct FpTraits;
template <> struct FpTraits<float> {
    static constexpr float eps = 1.1920928955078125e-07f;     // 2^-23 == numeric_limits<float>::epsilon()
    static constexpr float unit_roundoff = 5.9604644775390625e-08f;  // 2^-24
    static constexpr float sentinel = 1.0e18f;                 // (1e18)^2 * 12 < FLT_MAX
    static constexpr float big = 3.402823466e+38f;             // numeric_limits<float>::max()
    static constexpr float tiny = 1.0e-35f;
    static constexpr float xSentinel = 1.0e30f;                // expanded-screen value of empty slots
    static constexpr float minNormal = 1.17549435e-38f;        // numeric_limits<float>::min()
};
template <> struct FpTraits<double> {
    static constexpr double eps = 2.220446049250313080847e-16;  // 2^-52
    static constexpr double unit_roundoff = 1.1102230246251565404e-16;
    static constexpr double sentinel = 1.0e150;
    static constexpr double big = 1.7976931348623157e+308;
    static constexpr double tiny = 1.0e-300;
    static constexpr double xSentinel = 1.0e300;
    static constexpr double minNormal = 2.2250738585072014e-308;
};

// Per-segment static description (host-built, uploaded once per run).
struct SegDesc {
    int64_t pOff;    // first point (global index into the SoA point buffer)
    int32_t nP;      // points the segment evolves over (own, then neighbours)
    int32_t nOwn;    // points the segment chooses from (a prefix of the nP points)
    int64_t sOff;    // first chosen slot
    int32_t sCap;    // chosen capacity (multiple of kChosenTile)
    int32_t group;   // decimate group: segments of one group decimate together
};

// A unit of evolve-scan work: points [start, end) of segment seg (segment-relative).
struct WorkItem {
    int32_t seg;
    int32_t start;
    int32_t end;
    int32_t pad;
};

// Segments address their samples by position; with an index list (tiles, whose neighbour
// samples repeat other tiles' points) position p holds point pIndex[p] of the SoA buffer.
__device__ __forceinline__ int64_t point_of(const int* __restrict__ pIndex, int64_t pos) {
    return pIndex ? (int64_t)pIndex[pos] : pos;
}

// ---------------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ uint64_t lanes_below() { return __lanemask_lt(); }

// XCD-aware block order: the GPU hands consecutive workgroups of a launch to its 8 XCDs in
// turn, and each XCD has its own L2.  Block i of an n-block launch takes item xcd_item(i, n),
// a bijection of [0, n) under which the blocks of one XCD take one contiguous run of items,
// so neighbouring items' partial cache lines meet in the same L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_item(int i, int n) {
    const int q = n / kXcds, r = n - q * kXcds;
    const int x = i % kXcds, j = i / kXcds;
    return x * q + (x < r ? x : r) + j;
}

// Exact squared distance in the reference's operation order (no contraction).
template <typename T, int DIM>
__device__ __forceinline__ T exact_sqdist(const T (&p)[DIM], const T (&s)[DIM]) {
    T acc = T(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        T t = s[d] - p[d];
        T sq = t * t;
        acc = acc + sq;
    }
    return acc;
}

template <typename T>
__device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }

__device__ __forceinline__ float min3f(float a, float b, float c) { return fminf(fminf(a, b), c); }

// second-smallest update: with best <= second, returns the second smallest of
// {best, second, m}.
__device__ __forceinline__ float second_update(float best, float second, float m) {
    return __builtin_amdgcn_fmed3f(best, second, m);
}
__device__ __forceinline__ double second_update(double best, double second, double m) {
    return fmin(second, fmax(best, m));
}

}  // namespace rdhip

// Dimension dispatch: instantiates BODY with a constexpr int DIM for d in [1, 12].
#define RD_DISPATCH_DIM(d, ...)                                                      \
    switch (d) {                                                                     \
        case 1: { constexpr int DIM = 1; __VA_ARGS__; } break;                       \
        case 2: { constexpr int DIM = 2; __VA_ARGS__; } break;                       \
        case 3: { constexpr int DIM = 3; __VA_ARGS__; } break;                       \
        case 4: { constexpr int DIM = 4; __VA_ARGS__; } break;                       \
        case 5: { constexpr int DIM = 5; __VA_ARGS__; } break;                       \
        case 6: { constexpr int DIM = 6; __VA_ARGS__; } break;                       \
        case 7: { constexpr int DIM = 7; __VA_ARGS__; } break;                       \
        case 8: { constexpr int DIM = 8; __VA_ARGS__; } break;                       \
        case 9: { constexpr int DIM = 9; __VA_ARGS__; } break;                       \
        case 10: { constexpr int DIM = 10; __VA_ARGS__; } break;                     \
        case 11: { constexpr int DIM = 11; __VA_ARGS__; } break;                     \
        case 12: { constexpr int DIM = 12; __VA_ARGS__; } break;                     \
        default: break;                                                              \
    }
