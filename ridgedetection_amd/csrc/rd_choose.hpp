// rd_choose.hpp -- launchers for the greedy choose kernels (rd_choose.hip).
#pragma once

#include "rd_common.hpp"

namespace rdhip {

template <typename T>
struct ChooseArgs {
    const T* P;
    int64_t pStride;
    const int* pIndex;        // optional position -> point map
    T* S;
    int64_t sStride;
    const SegDesc* segs;
    int numSegs;
    int* segK;
    const T* segRSqr;
    unsigned char* covered;   // per sample: covered by the chosen set at its chunk's start
    int* chosenIdx;           // per chosen slot: segment-relative index of the chosen sample
    int* overflow;
    const int* segLarge;      // per segment: 1 = chunked path (filter + block), 0 = one block
};

// Chunk schedule of the large-segment path: chunk c covers samples
// [choose_chunk_start(c), choose_chunk_start(c+1)).  Chunk 0 is the sequential start.
__host__ __device__ inline int choose_chunk_start(int c) {
    if (c <= 0) return 0;
    if (c == 1) return 2048;
    if (c == 2) return 8192;
    if (c == 3) return 32768;
    long long v = 32768LL + (long long)(c - 3) * 65536LL;
    return v > 0x7fffffffLL ? 0x7fffffff : (int)v;
}

template <typename T>
hipError_t launch_choose_filter(const ChooseArgs<T>& a, int dim, const WorkItem* items, int numItems,
                                hipStream_t st);
template <typename T>
hipError_t launch_choose_block(const ChooseArgs<T>& a, int dim, int chunk, int useFlags, hipStream_t st);

}  // namespace rdhip
