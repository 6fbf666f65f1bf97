// rd_quality.hpp -- distances of a point set to an ideal pattern (rd_quality.hip).
#pragma once

#include "rd_common.hpp"

#include <vector>

namespace rdhip {

// Host ROW_MAJOR inputs.  For every point of A, the distance to the nearest point of B.
template <typename T>
std::vector<T> set_distances(const T* A, int64_t nA, const T* B, int64_t nB, int dim, hipStream_t st);
// max(setDistance(A, B), setDistance(B, A)); -1 for an empty A.
template <typename T>
T hausdorff_distance(const T* A, int64_t nA, const T* B, int64_t nB, int dim, hipStream_t st);
// (median, average, min, max) of the distances from A to B.
template <typename T>
void distance_stats(const T* A, int64_t nA, const T* B, int64_t nB, int dim, hipStream_t st, T* out4);

}  // namespace rdhip
