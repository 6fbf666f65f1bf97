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
};

struct DecWorkspace {
    int* partials = nullptr;   // scan partials for max(numSegs, maxElems + 1)
    int64_t* nbrTotal = nullptr;
    int* nbr = nullptr;
    int64_t nbrCap = 0;
};

// Decimates every group; chosenIdx (optional) is compacted alongside the coordinates and
// keepOut (optional, per slot) receives 1 for kept / 0 for removed points.  Synchronises
// the stream once (to size the neighbour lists).
template <typename T>
hipError_t decimate_groups(DecArgs& a, T* S, int64_t sStride, int dim, const T* groupRSqr, int* chosenIdx,
                           unsigned char* keepOut, int64_t maxElems, DecWorkspace& ws, hipStream_t st);

}  // namespace rdhip
