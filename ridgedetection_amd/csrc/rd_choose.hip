// rd_choose.hip -- greedy choose (rd/cpu/brute_force/choose.hpp:36-70) for a batch of segments.
//
// Semantics: sample 0 is chosen; sample i is chosen iff no already chosen sample lies at
// squared distance <= r*r (exact reference distance).  The decision depends on every
// earlier decision, so the samples are processed in order in sub-chunks of 256:
//   * every thread tests its sample against the chosen points known when the sub-chunk
//     starts (uniform loop: the chosen coordinates are scalar loads);
//   * the survivors are compacted in order (wave ballots + a block prefix) and resolved
//     one after another against the points chosen inside the sub-chunk, which are kept
//     in LDS; a survivor not covered by any of them is appended.
// For large segments the first test runs on the whole GPU (choose_filter_kernel) one
// chunk at a time, against the chosen set at the start of the chunk; the block kernel
// then resolves only that chunk.  Output per segment: K, the chosen coordinates in the
// segment's slots and the chosen samples' indices (segment-relative, ascending).
#include "rd_choose.hpp"

namespace rdhip {

namespace {
constexpr int kThreads = 256;
}

template <typename T, int DIM>
__device__ __forceinline__ bool covered_by(const T (&p)[DIM], const T* __restrict__ Sb, int64_t ss, int k0,
                                           int k1, T rSqr) {
    for (int k = k0; k < k1; ++k) {
        T s[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) s[d] = Sb[d * ss + k];
        if (exact_sqdist<T, DIM>(p, s) <= rSqr) return true;
    }
    return false;
}

template <typename T, int DIM>
__global__ __launch_bounds__(kThreads) void choose_filter_kernel(ChooseArgs<T> a, const WorkItem* __restrict__ items) {
    const WorkItem it = items[blockIdx.x];
    const SegDesc sd = a.segs[it.seg];
    const int K = a.segK[it.seg];
    const T rSqr = a.segRSqr[it.seg];
    const T* __restrict__ Sb = a.S + sd.sOff;
    for (int i = it.start + threadIdx.x; i < it.end; i += kThreads) {
        T p[DIM];
        const int64_t src = point_of(a.pIndex, sd.pOff + i);
#pragma unroll
        for (int d = 0; d < DIM; ++d) p[d] = a.P[d * a.pStride + src];
        a.covered[sd.pOff + i] = covered_by<T, DIM>(p, Sb, a.sStride, 0, K, rSqr) ? 1 : 0;
    }
}

template <typename T, int DIM>
__global__ __launch_bounds__(kThreads) void choose_block_kernel(ChooseArgs<T> a, int chunk, int useFlags) {
    __shared__ T newS[kThreads][DIM];
    __shared__ int surv[kThreads];
    __shared__ int waveCnt[kThreads / kWave];

    const int seg = blockIdx.x;
    if (seg >= a.numSegs) return;
    if (a.segLarge && (a.segLarge[seg] != 0) != (chunk >= 0)) return;
    const SegDesc sd = a.segs[seg];
    const int n = sd.nOwn;
    const T rSqr = a.segRSqr[seg];
    const int64_t ss = a.sStride;
    T* __restrict__ Sb = a.S + sd.sOff;
    const int64_t ps = a.pStride;
    int* __restrict__ cidx = a.chosenIdx + sd.sOff;
    auto coord = [&](int i, int d) { return a.P[d * ps + point_of(a.pIndex, sd.pOff + i)]; };

    int from, to;
    if (chunk < 0) {
        from = 0;
        to = n;
    } else {
        from = choose_chunk_start(chunk);
        to = min(choose_chunk_start(chunk + 1), n);
    }
    if (from >= to) {
        if (chunk <= 0 && n == 0 && threadIdx.x == 0) a.segK[seg] = 0;
        return;
    }
    int K;
    if (from == 0) {
        // sample 0 is always chosen
        if (threadIdx.x < DIM) Sb[threadIdx.x * ss] = coord(0, threadIdx.x);
        if (threadIdx.x == 0) cidx[0] = 0;
        K = 1;
        from = 1;
        __syncthreads();
    } else {
        K = a.segK[seg];
    }
    const int kChunkStart = K;   // chosen set the filter pass (if any) tested against
    const int cap = sd.sCap;
    const int wave = threadIdx.x / kWave;

    for (int c0 = from; c0 < to; c0 += kThreads) {
        const int i = c0 + threadIdx.x;
        T p[DIM];
        bool live = i < to;
#pragma unroll
        for (int d = 0; d < DIM; ++d) p[d] = live ? coord(i, d) : T(0);
        if (live) {
            if (useFlags && a.covered[sd.pOff + i]) live = false;
            else if (covered_by<T, DIM>(p, Sb, ss, useFlags ? kChunkStart : 0, K, rSqr)) live = false;
        }
        // compact survivors in sample order
        const uint64_t m = __ballot(live);
        if (lane_id() == 0) waveCnt[wave] = __popcll(m);
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kThreads / kWave; ++w) {
            before += (w < wave) ? waveCnt[w] : 0;
            total += waveCnt[w];
        }
        if (live) surv[before + __popcll(m & lanes_below())] = threadIdx.x;
        __syncthreads();
        // resolve survivors one by one against the points chosen in this sub-chunk
        int added = 0;
        for (int q = 0; q < total; ++q) {
            const int t = surv[q];
            bool cov = false;
            if (threadIdx.x < added) {
                T pq[DIM], s[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) {
                    pq[d] = coord(c0 + t, d);
                    s[d] = newS[threadIdx.x][d];
                }
                cov = exact_sqdist<T, DIM>(pq, s) <= rSqr;
            }
            if (__syncthreads_or(cov)) continue;
            if (threadIdx.x == t) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) newS[added][d] = p[d];
                if (K + added < cap) cidx[K + added] = i;
            }
            ++added;
            __syncthreads();
        }
        // append this sub-chunk's chosen points
        if (threadIdx.x < added) {
            const int k = K + threadIdx.x;
            if (k < cap) {
#pragma unroll
                for (int d = 0; d < DIM; ++d) Sb[d * ss + k] = newS[threadIdx.x][d];
            }
        }
        __syncthreads();
        K += added;
        if (K > cap) {
            if (threadIdx.x == 0) atomicOr(a.overflow, 1);
            K = cap;
            break;
        }
    }
    if (threadIdx.x == 0) a.segK[seg] = K;
}

template <typename T>
hipError_t launch_choose_filter(const ChooseArgs<T>& a, int dim, const WorkItem* items, int numItems,
                                hipStream_t st) {
    if (numItems <= 0) return hipSuccess;
    bool ok = false;
    RD_DISPATCH_DIM(dim, {
        hipLaunchKernelGGL((choose_filter_kernel<T, DIM>), dim3(numItems), dim3(kThreads), 0, st, a, items);
        ok = true;
    });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_choose_block(const ChooseArgs<T>& a, int dim, int chunk, int useFlags, hipStream_t st) {
    if (a.numSegs <= 0) return hipSuccess;
    bool ok = false;
    RD_DISPATCH_DIM(dim, {
        hipLaunchKernelGGL((choose_block_kernel<T, DIM>), dim3(a.numSegs), dim3(kThreads), 0, st, a, chunk,
                           useFlags);
        ok = true;
    });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template hipError_t launch_choose_filter<float>(const ChooseArgs<float>&, int, const WorkItem*, int, hipStream_t);
template hipError_t launch_choose_filter<double>(const ChooseArgs<double>&, int, const WorkItem*, int, hipStream_t);
template hipError_t launch_choose_block<float>(const ChooseArgs<float>&, int, int, int, hipStream_t);
template hipError_t launch_choose_block<double>(const ChooseArgs<double>&, int, int, int, hipStream_t);

}  // namespace rdhip
