// rd_chain.hpp -- member ordering for the evolve iteration of one segment with up to
// kChainMaxKeys chosen points (a whole cloud) without a separate histogram pass: the scan
// counts labels per block, one kernel turns the counts into per-key prefixes and one scatter
// writes the member coordinates in chain order.  Kernels in rd_chain.hip.
#pragma once

#include "rd_evolve.hpp"

namespace rdhip {

// chosen points per segment the path handles (the scan's per-block label histogram lives in LDS)
constexpr int kChainMaxKeys = 2048;
// dimensions the path is built for (the scatter keeps DIM coordinates per sample in registers)
constexpr int kChainMaxDim = 4;

// Chains are laid out in key order, chain k in [chainOff[k], chainEnd[k]) with every chain END
// aligned to kChainAlign<T> values (16 bytes).  Gaps between chains are never read.
template <typename T>
constexpr int kChainAlign = 16 / (int)sizeof(T);

// Stride of the member coordinate rows for n samples and at most kCap keys (the alignment
// gaps add fewer than 4 values per key).
inline int64_t chain_member_stride(int64_t n, int64_t kCap) {
    const int64_t m = n + 4 * kCap + 64;
    return (m + 63) / 64 * 64;
}

template <typename T>
struct ChainArgs {
    // samples of the one segment
    const T* P;
    int64_t pStride;
    const int* pIndex;
    const WorkItem* items;   // the scan's items, in sample order; item b's samples are block b's
    const SegDesc* segs;
    const int* segK;
    const int* labels;       // per segment position: chosen slot, or -1
    // hist[k * histStride + b]: the scan's per-block label counts; key_prefix turns every
    // row into its exclusive prefix over blocks
    int* hist;
    int64_t histStride;
    int* totals;             // members per key
    T* sortedP;              // member coordinates, SoA rows of mStride, chain layout above
    int64_t mStride;
    const int* allDone;
    int keyBound;            // keys handled: k < min(segK[0], keyBound) (the buffers' capacity)
    // written by the scatter's block 0 for the ordered sums (ShiftArgs of rd_evolve.hpp):
    // chain ranges, the wave-summed chain list over longEntries multiples of 2^longShift, and
    // the end of the last chain
    int* chainOff;
    int* chainEnd;
    int* longList;
    int longShift;
    int longEntries;
    int64_t* mCount;
};

// Launches the per-key prefix and the member scatter, for a scan that wrote labels and
// blockHist over numItems items of itemSize samples.  kb (== a.keyBound) bounds the number of
// keys; the device count segK[0] is authoritative below it.
template <typename T>
hipError_t launch_chain_order(const ChainArgs<T>& a, int dim, int numItems, int itemSize, int kb, hipStream_t st);

// True when launch_chain_order supports this shape.
bool chain_iteration_supported(int dim, int itemSize, int64_t kb);

}  // namespace rdhip
