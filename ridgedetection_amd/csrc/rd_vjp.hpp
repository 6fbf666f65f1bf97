// rd_vjp.hpp -- backward (vector-Jacobian) kernels of the evolve iterations with frozen
// memberships (rd_vjp.hip).
#pragma once

#include "rd_common.hpp"

namespace rdhip {

// One logged run of `steps` iterations, newest last.  Per step u:
//   labels[u * lStride + i]  chosen slot of position i, or -1
//   counts[u * cStride + k]  members of slot k
//   moved[(u * dim + d) * cStride + k]   1 when coordinate d of slot k moved
template <typename T>
struct VjpArgs {
    const int* labels;
    int64_t lStride;
    const int* counts;
    const unsigned char* moved;
    int64_t cStride;
    int steps;
    int K;              // live slots (0 .. K-1)
    T* G;               // dL/dS, SoA [dim][gStride]; dL/dS_t on return of the coefficient pass
    int64_t gStride;
    T* A;               // coefficients, [steps][dim][aStride]
    int64_t aStride;
    T* gradP;           // dL/dP, SoA [dim][gpStride], accumulated over positions 0 .. n-1
    int64_t gpStride;
    int64_t n;          // positions = points (one segment, no index map)
    int accumulate;     // 0: the first logged run overwrites gradP
};

// Coefficient pass (slots, newest step first), then the member pass (positions).
template <typename T>
hipError_t launch_vjp_run(const VjpArgs<T>& a, int dim, hipStream_t st);

// gradP[:, idx[k]] += G[:, k] for k < K (S_0 drawn from the samples idx; distinct indices).
template <typename T>
hipError_t launch_vjp_seed(const T* G, int64_t gStride, const int* idx, int K, int dim, T* gradP, int64_t gpStride,
                           hipStream_t st);

}  // namespace rdhip
