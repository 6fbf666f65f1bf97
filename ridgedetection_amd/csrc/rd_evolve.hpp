// rd_evolve.hpp -- launchers for the evolve iteration kernels (rd_evolve.hip).
#pragma once

#include "rd_common.hpp"

#include <algorithm>

namespace rdhip {

constexpr int kScanThreads = 256;

template <typename T>
struct ScanArgs {
    const T* P;
    int64_t pStride;
    const int* pIndex;      // optional position -> point map
    const T* S;
    int64_t sStride;
    const WorkItem* items;
    const SegDesc* segs;
    const int* segK;
    const int* segDone;
    const T* segRSqr;
    const int* allDone;
    int* labels;   // per sample: chosen slot, or -1 when outside every r-ball
    int* queue;    // work-queue head of the persistent form (reset every iteration)
    // compact form (batches of segments): item it writes (label, position) pairs to
    // outKeys/outPos[it.pad + (i - it.start)] instead of labels[]; the persistent loop
    // stops at *numItemsDev, so the active item list can shrink between graph replays
    int* outKeys;
    int* outPos;
    const int* numItemsDev;
    // optional (compact form): the key written for slot s is denseOf[s], its index among
    // the live slots of the evolve call, so that the member sort spans only those
    const int* denseOf;
    // optional (one block per item): every block also writes the histogram of its item's
    // labels below min(K, histKeyCap[seg]) into its segment's table,
    // blockHist[histSegOff[seg] + k * nb + start / histItem] with nb = ceil(nP / histItem)
    // (rd_chain.hpp), and blocks take the items in use in the XCD-aware order of xcd_item()
    int* blockHist;
    const int64_t* histSegOff;
    const int* histKeyCap;
    int histItem;
    // optional: the expanded-form screen (rd_nearest.hpp) over the table of expand_chosen
    // (rows d < dim: s - o, row dim: |s - o|^2, stride sStride) and the per-segment
    // centre and bound segX[seg * kScreenStride + d], segX[seg * kScreenStride + kMaxDim]
    const T* xTable;
    const T* segX;
    int mfma;   // 1: the expanded screen on the matrix cores (FP32; nearest_labels_mfma)
};

constexpr int kScreenStride = kMaxDim + 1;

// Per-segment centre o (midpoint of the live chosen points' bounding box), the expanded table
// rows of every slot below the tile-rounded live count (slots past it: s' = 0, b = xSentinel)
// and the largest b over the live slots; one block per segment, skipped for converged
// segments and after the all-done flag is set.
template <typename T>
hipError_t launch_expand_chosen(const T* S, int64_t sStride, const SegDesc* segs, const int* segK,
                                const int* segDone, int numSegs, int dim, T* xTable, T* segX, const int* allDone,
                                hipStream_t st);

template <typename T>
struct ShiftArgs {
    const T* P;
    int64_t pStride;
    const int* pIndex;      // optional position -> point map
    T* S;
    int64_t sStride;
    int64_t totalSlots;
    const int* slotSeg;     // owning segment of every chosen slot (-1: padding)
    const SegDesc* segs;
    const int* segK;
    const int* segDone;
    int* segMoved;
    const int* members;     // sample indices sorted by (label, sample)
    const int64_t* mPtr;    // number of members (device)
    T* sortedP;             // SoA member coordinates in chain order (scratch)
    int64_t mStride;
    const int* chainOff;
    const int* chainEnd;
    const int* allDone;
    int* outCounts;         // optional: members per slot
    T* outSums;             // optional: SoA coordinate sums per slot
    unsigned char* outMoved;   // optional: 1 where coordinate d of a slot moved, at [d * movedStride + slot]
    int64_t movedStride;
    // chains summed by a whole wave instead of a 16-lane row: those holding a sorted
    // position that is a multiple of L = 2^longShift (so every chain of at least L members);
    // longList[m] is the chain whose first such position is m L, else -1 (chain_bounds)
    const int* longList;
    int longShift;
    // optional: the slots the rows visit (row r: coordinate r / totalSlots, slot
    // rowSlots[r % totalSlots]); null: every slot below totalSlots
    const int* rowSlots;
    // chains are indexed by the dense key (row r's chain is r, a listed chain's slot is
    // rowSlots[key]) instead of the slot
    int denseChains;
    // 1: tolerance mode -- every chain is summed in parallel in a wider type (double for
    // float data, double-double for double data) in a fixed order, instead of serially in the
    // reference's sample order; the mean is that sum over the count, rounded once
    int tolerance;
    // optional (member layouts of rd_chain.hip): the wave list by length instead of by position
    // -- every chain of at least longMin members took an index from *longCount (longIdx[chain]);
    // the first longCap of them are summed by waves (longList[index] = chain), the rest by rows
    const int* longCount;
    const int* longIdx;
    int longMin;
    int longCap;
};

// Mean of a chain and the shift rule of rd/cpu/brute_force/evolve.hpp:129-179 for one
// coordinate; optional per-slot outputs.
template <typename T>
__device__ __forceinline__ void shift_coordinate(const ShiftArgs<T>& a, int d, int64_t slot, int seg, int cnt,
                                                 T sum) {
    if (a.outCounts && d == 0) a.outCounts[slot] = cnt;
    if (a.outSums) a.outSums[d * a.sStride + slot] = sum;
    bool moved = false;
    if (cnt != 0) {
        const T fc = (T)cnt;
        constexpr T eps = FpTraits<T>::eps;
        T mc = sum / fc;
        T thr = T(2) * fabs(mc);
        thr = thr * fc;
        thr = thr * eps;
        T& s = a.S[d * a.sStride + slot];
        if (fabs(mc - s) > thr) {
            s = mc;
            moved = true;
        }
    }
    if (a.outMoved) a.outMoved[d * a.movedStride + slot] = moved ? 1 : 0;
    if (moved) a.segMoved[seg] = 1;
}

// Per-segment iteration count and convergence after an iteration, and the batch's all-done
// flag, by one block of any size.
__device__ __forceinline__ void finalize_segments(int numSegs, int* __restrict__ segDone, int* __restrict__ segMoved,
                                                  int* __restrict__ segIter, int maxIter, int* __restrict__ allDone,
                                                  int* __restrict__ totalIters) {
    __shared__ int active;
    if (threadIdx.x == 0) active = 0;
    __syncthreads();
    int mine = 0;
    for (int s = threadIdx.x; s < numSegs; s += blockDim.x) {
        if (segDone[s]) continue;
        int it = segIter[s] + 1;
        segIter[s] = it;
        bool done = (segMoved[s] == 0) || (maxIter > 0 && it >= maxIter);
        segMoved[s] = 0;
        if (done) segDone[s] = 1; else ++mine;
    }
    if (mine) atomicAdd(&active, mine);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (active == 0) *allDone = 1;
        *totalIters += 1;
    }
}

// L of the wave-summed chains for up to maxM sorted members: the smallest power of two
// >= 2^kLongMinShift with at most kLongWaves (coordinate, list entry) waves.
constexpr int kLongMinShift = 10;
constexpr int kLongWaves = 2048;
inline int long_chain_shift(int dim, int64_t maxM) {
    int sh = kLongMinShift;
    while ((int64_t)dim * ((maxM + (int64_t(1) << sh) - 1) >> sh) > kLongWaves) ++sh;
    return sh;
}
inline int64_t long_list_entries(int64_t maxM, int shift) { return std::max<int64_t>((maxM >> shift) + 1, 1); }

// persistentBlocks > 0: that many persistent blocks (of kScanThreads) pull the work items
// from a device queue; 0: one block of `threads` (128, 256 or 512) per work item, each item
// holding threads x ppt samples.
template <typename T>
hipError_t launch_evolve_scan(const ScanArgs<T>& a, int dim, int numItems, int ppt, int threads,
                              int persistentBlocks, hipStream_t st);
// Sums and shifts (after gathering the members' coordinates into chain order when
// `gather`; otherwise a.sortedP already holds them); slots [0, a.totalSlots) are visited
// (pass the live bound, not the capacity).
template <typename T>
hipError_t launch_chain_sum_shift(const ShiftArgs<T>& a, int dim, int64_t maxM, bool gather, hipStream_t st);
// longList (optional): the wave-summed chain list of ShiftArgs for L = 2^longShift
hipError_t launch_chain_bounds(const int* keys, const int64_t* dM, int64_t maxM, int* chainOff,
                               int* chainEnd, int* longList, int longShift, const int* allDone,
                               hipStream_t st);
hipError_t launch_evolve_finalize(int numSegs, int* segDone, int* segMoved, int* segIter, int maxIter,
                                  int* allDone, int* totalIters, hipStream_t st);
template <typename T>
hipError_t launch_fill_sentinel(T* S, int64_t sStride, int dim, const SegDesc* segs, const int* segK,
                                int numSegs, hipStream_t st);

}  // namespace rdhip
