// rd_local.hpp -- evolve of small segments, one workgroup per segment (rd_local.hip).
#pragma once

#include "rd_common.hpp"

namespace rdhip {

constexpr int kLocalThreads = 512;
constexpr int kLocalWaves = kLocalThreads / kWave;
constexpr int kLocalMaxP = 65536;     // samples (positions) of a segment taken by the local path
constexpr int kLocalChunk = 8192;     // samples per chunk inside an iteration (uint16 ranks)
constexpr int kLocalMaxK = 1024;      // chosen points of a segment
constexpr size_t kLocalLdsBudget = 128 * 1024;

template <typename T>
struct LocalArgs {
    const T* P;
    int64_t pStride;
    const int* pIndex;      // optional position -> point map
    T* S;
    int64_t sStride;
    const SegDesc* segs;
    const int* segK;
    const T* segRSqr;
    const int* list;        // segments to evolve, one workgroup each
    int count;
    int kStride;            // LDS stride of the chosen coordinates (>= every listed K, multiple of 16)
    int pCap;               // samples per chunk (<= kLocalChunk)
    int maxIter;            // 0: until convergence
    int guard;              // an uncapped evolve stops here and raises *error
    int* segIter;           // iterations run, per segment
    int* maxIters;          // max over the listed segments (atomicMax)
    int* error;
    // one-iteration form (launch_segment_sums): labels of every position from the
    // batch-wide scan, per-segment done / moved flags, the batch's all-done flag
    const int* labelsIn;
    const int* segDone;
    int* segMoved;
    const int* allDone;
};

// LDS bytes the local evolve needs for chosen stride kStride and chunks of `chunk` samples.
size_t local_lds_bytes(int dim, size_t valueBytes, int kStride, int chunk);

// Runs every listed segment's evolve to convergence (or maxIter iterations): the
// reference's evolve loop (rd/cpu/brute_force/evolve.hpp:237-459) with every iteration's
// labels, member lists, ordered sums and shift inside one workgroup.
template <typename T>
hipError_t launch_local_evolve(const LocalArgs<T>& a, int dim, hipStream_t st);

// One evolve iteration for segments 0 .. numSegs-1 (one workgroup each, done segments
// skipped) from the labels of the batch-wide scan: counting sort, ordered sums and the
// shift of rd/cpu/brute_force/evolve.hpp:32-77, 129-179, in LDS.
template <typename T>
hipError_t launch_segment_sums(const LocalArgs<T>& a, int dim, int numSegs, hipStream_t st);

}  // namespace rdhip
