// rd_decimate.hpp -- order-exact decimation over groups of segments (rd_decimate.hip).
#pragma once

#include "rd_common.hpp"

#include <algorithm>

namespace rdhip {

struct DecArgs {
    const SegDesc* segs;
    int numSegs;
    int* segK;                 // in: live counts; out: counts after decimation
    const int* groupSegBegin;  // group g = segments [begin, end), consecutive
    const int* groupSegEnd;
    int numGroups;
    int noCopyRule;            // 0: decimate (stop below 3), 1: decimateNoCopy (keep)
    int* elemBase;             // per segment: first element of the segment
    int64_t* totalElems;
    int* groupOff;
    int* groupCnt;
    int* groupLeft;
    int* elemSlot;
    int* elemGroup;
    int* c1;
    int* c2;
    int* rowOff;               // maxElems + 1 entries
    int* nbr;
    int* alive;
    const int* segSkip;        // optional, per segment: leave this segment untouched
    // groups of at least kDecRoundsMin elements decide in dependency rounds (set by
    // decimate_groups from its workspace)
    int* groupMode;            // kDecSerial / kDecRounds / kDecResume / kDecDone
    int* groupAlive;           // live elements at the start of the current pass
    int* groupRemoved;
    int* groupApply;
    int* state;                // per element in the current pass: 0 undecided, 1 kept, 2 removed
    int* prog;                 // next CSR entry an undecided element examines
    int* rc1;                  // removed earlier neighbours within r2 / 2 r2 found so far
    int* rc2;
    int* parFlags;             // [0] an element is still undecided, [1] groups in rounds, [2] groups to resume
};

constexpr int kDecRoundsMin = 2048;
enum : int { kDecSerial = 0, kDecRounds = 1, kDecResume = 2, kDecDone = 3 };

struct DecWorkspace {
    int* partials = nullptr;   // scan partials for max(numSegs, maxElems + 1)
    int64_t* nbrTotal = nullptr;
    int* nbr = nullptr;
    int64_t nbrCap = 0;
    int* par = nullptr;        // dependency-round state (DecArgs::groupMode ... parFlags)
    int64_t parCap = 0;
    void* boxes = nullptr;     // bounding boxes of element chunks (value type of the call)
    int64_t boxCap = 0;        // bytes
};

// Decimates every group; chosenIdx (optional) is compacted alongside the coordinates and
// keepOut (optional, per slot) receives 1 for kept / 0 for removed points.  Synchronises
// the stream once (to size the neighbour lists).
template <typename T>
hipError_t decimate_groups(DecArgs& a, T* S, int64_t sStride, int dim, const T* groupRSqr, int* chosenIdx,
                           unsigned char* keepOut, int64_t maxElems, DecWorkspace& ws, hipStream_t st);

}  // namespace rdhip
