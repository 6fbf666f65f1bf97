// rd_chain.hpp -- member ordering for evolve iterations without a sort, for segments (whole
// clouds of a batch, tiles) with up to kChainMaxKeys chosen points each: the scan counts labels
// per block, one kernel turns the counts into per-key prefixes over the segment's blocks and
// one scatter writes the member coordinates in chain order.  Kernels in rd_chain.hip.
#pragma once

#include "rd_evolve.hpp"

namespace rdhip {

// chosen points per segment the path handles (the scan's per-block label histogram lives in LDS)
constexpr int kChainMaxKeys = 2048;
// dimensions the path is built for (the scatter keeps DIM coordinates per sample in registers)
constexpr int kChainMaxDim = kMaxDim;

// Chains are laid out in key order, chain k in [chainOff[k], chainEnd[k]) with every chain END
// aligned to kChainAlign<T> values (16 bytes).  Gaps between chains are never read.
template <typename T>
constexpr int kChainAlign = 16 / (int)sizeof(T);

// Positions segment s's chains may take: its samples plus the alignment gaps (fewer than 4
// values per key), rounded to 64.
inline int64_t chain_region(int64_t nP, int64_t keys) { return (nP + 4 * keys + 64 + 63) / 64 * 64; }

// Per-segment layout (host-built, device arrays): segment s orders keys k < min(segK[s],
// segKeyCap[s]); its per-block counts are the table hist[segHistOff[s] + k * nb_s + col] with
// nb_s = ceil(nP_s / itemSize) blocks (col = item start / itemSize); its per-key totals are
// totals[segRowOff[s] + k]; its chains take the positions [segMemberOff[s], segMemberOff[s + 1]).
template <typename T>
struct ChainArgs {
    const T* P;
    int64_t pStride;
    const int* pIndex;
    const WorkItem* items;      // the scan's items (any order); item b covers block b
    const int* numItemsDev;     // optional: the items in use (else one per block)
    const SegDesc* segs;
    const int* segK;
    const int* segDone;
    int numSegs;
    const int* labels;          // per position: chosen slot, or -1
    int* hist;
    const int64_t* segHistOff;
    const int* segKeyCap;
    const int* segRowOff;       // numSegs + 1 entries
    int itemSize;
    int totalRows;
    int* totals;
    T* sortedP;                 // member coordinates, SoA rows of mStride
    int64_t mStride;
    const int64_t* segMemberOff;   // numSegs + 1 entries
    const int* allDone;
    // written by each segment's first block for the ordered sums (ShiftArgs of rd_evolve.hpp):
    // chain ranges per slot, and for every chain of at least longMin members an index from
    // *longCount (reset by the prefix pass) into longIdx[slot], the first longCap of them listed
    // in longList for the wave-summed path
    int* chainOff;
    int* chainEnd;
    int* longList;
    int* longCount;
    int* longIdx;
    int longMin;
    int longCap;
};

// Launches the per-key prefix and the member scatter after a histogram scan over numItems
// blocks (the items in use may be fewer: a.numItemsDev); keyCapMax bounds segKeyCap.
template <typename T>
hipError_t launch_chain_order(const ChainArgs<T>& a, int dim, int numItems, int keyCapMax, hipStream_t st);

// True when launch_chain_order supports this shape.
bool chain_iteration_supported(int dim, int itemSize, int64_t keyCapMax);

}  // namespace rdhip
