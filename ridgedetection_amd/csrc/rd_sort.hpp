// rd_sort.hpp -- stable device sort of int32 (key, value) pairs by key.
#pragma once

#include "rd_common.hpp"

namespace rdhip {

// Workspace needed by stable_sort_by_key for up to maxN pairs with keys in [0, numKeys).
size_t sort_workspace_bytes(int64_t maxN, int numKeys);

// True when a sort over keys in [0, numKeys) runs in one pass (numKeys <= 2^11); its scanned
// digit histogram then holds every key's output range.
bool sort_single_pass(int numKeys);
// After such a one-pass stable_sort_by_key (same workspace, maxN, numKeys; *dM pairs kept):
// off[k] / end[k] = the range of key k in the sorted output for every k < numKeys, and,
// when longList is given, longList[m] = the key whose range first holds position m 2^longShift
// (-1 for the other multiples below *dM).  Returns hipErrorInvalidValue for a multi-pass sort.
hipError_t launch_key_ranges_from_sort(const void* workspace, int64_t maxN, int numKeys, const int64_t* dM,
                                       int* off, int* end, int* longList, int longShift, const int* skip,
                                       hipStream_t st);

// Sorts the *dN pairs (keysIn[i], valsIn ? valsIn[i] : i) by key, keeping input order
// among equal keys; pairs with a negative key are dropped.  The number of kept pairs is
// written to *dM.  All counts live on the device, so the call is stream-ordered and
// capturable.  maxN bounds *dN and sizes the grids.  When `skip` is not null and *skip
// is set when the kernels run, they return at once (used by converged evolve loops).
hipError_t stable_sort_by_key(const int* keysIn, const int* valsIn, const int64_t* dN, int64_t maxN,
                              int numKeys, int* keysOut, int* valsOut, int64_t* dM, void* workspace,
                              hipStream_t stream, const int* skip = nullptr);

// Optional companion output of the last pass: for every kept pair, in sorted order, the
// coordinates of point point_of(pIndex, value) of the SoA cloud P (dim x pStride) go to
// out[d * outStride + rank] instead of writing the value itself (valsOut is not written).
template <typename T>
struct SortGather {
    const T* P;
    int64_t pStride;
    const int* pIndex;
    T* out;
    int64_t outStride;
    int dim;
};
template <typename T>
hipError_t stable_sort_by_key(const int* keysIn, const int* valsIn, const int64_t* dN, int64_t maxN,
                              int numKeys, int* keysOut, int* valsOut, int64_t* dM, void* workspace,
                              hipStream_t stream, const int* skip, const SortGather<T>& gather);

// Exclusive prefix sum of a[0..L) in place; `partials` needs scan_partials_count(L) ints;
// the grand total is written to *total when total is not null.
int64_t scan_partials_count(int64_t L);
hipError_t exclusive_scan_inplace(int* a, int64_t L, int* partials, int64_t* total, hipStream_t stream);

}  // namespace rdhip
