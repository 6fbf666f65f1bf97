// rd_evolve.hip -- one evolve iteration for a batch of segments, on gfx950.
//
// Reference semantics (one iteration, rd/cpu/brute_force/evolve.hpp:32-77, 129-179):
//   for every sample n (in order): k* = first argmin_k d2(S_k, P_n); if d2 <= r*r then
//   count[k*]++ and sum[k*] += P_n; then for every chosen k and coordinate d with
//   count > 0: mc = sum / count; if |mc - S| > ((2|mc|) count) eps: S = mc, flag = 1.
//
// Device pipeline per iteration:
//   1. evolve_scan  -- nearest chosen point of every sample.  The K chosen points of a
//      segment stream through SGPRs (uniform scalar loads) in tiles of 16; every lane
//      keeps 4 samples.  A screening distance (1 mul + DIM-1 fma) with a v_min3 tree
//      tracks, per sample, the best tile-minimum, the tile holding it and the second best
//      tile-minimum.  The epilogue re-scans the winning tile for the winner and the
//      in-tile runner-up, then decides exactly: the screening value differs from the
//      reference distance by at most delta = 4(DIM+1)u relative, so if the runner-up
//      exceeds best(1+3 delta) the screened winner is the reference argmin and only its
//      exact distance is computed (for the r-test); otherwise (near ties) the lane falls
//      back to the exact reference scan with strict-< first-index ties.  No MFMA: the
//      d <= 12 inner product feeds an argmin.
//   2. stable sort of the sample indices by label (rd_sort.hip): the members of every
//      chosen point in sample order.
//   3. chain_bounds -- start/end of every chosen point's member run.
//   4. gather_members + chain_sum_shift -- the members' coordinates are copied into chain
//      order; then one wave per chosen point reads 64 members per step (coalesced) and
//      folds them into the running sum lane by lane, i.e. in sample order (the reference's
//      serial summation, so sums are bit-identical), applies the shift rule and raises
//      the segment's "moved" flag.
//   5. evolve_finalize -- per-segment iteration count and convergence; global all-done.
#include "rd_evolve.hpp"

namespace rdhip {

// ---------------------------------------------------------------------------------
// 1. nearest chosen point
// ---------------------------------------------------------------------------------
template <typename T, int DIM>
__device__ __forceinline__ T screen_dist(const T (&p)[DIM], const T* s /* s[d] */) {
    T t0 = s[0] - p[0];
    T a = t0 * t0;
#pragma unroll
    for (int d = 1; d < DIM; ++d) {
        T t = s[d] - p[d];
        a = fma(t, t, a);
    }
    return a;
}

template <typename T>
__device__ __forceinline__ T tile_min(T m, T a, T b);
template <>
__device__ __forceinline__ float tile_min<float>(float m, float a, float b) { return min3f(m, a, b); }
template <>
__device__ __forceinline__ double tile_min<double>(double m, double a, double b) {
    return fmin(fmin(m, a), b);
}

template <typename T, int DIM, int PPT>
__device__ __forceinline__ void scan_item(const ScanArgs<T>& a, const WorkItem it) {
    const SegDesc sd = a.segs[it.seg];
    const int tid = threadIdx.x;
    int* __restrict__ lab = a.labels + sd.pOff;
    // compact output: slot of sample i is it.pad + (i - it.start)
    int* __restrict__ ck = a.outKeys ? a.outKeys + it.pad - it.start : nullptr;
    int* __restrict__ cp = a.outKeys ? a.outPos + it.pad - it.start : nullptr;
    auto put = [&](int i, int label) {
        if (ck) {
            ck[i] = label;
            cp[i] = (int)(sd.pOff + i);
        } else {
            lab[i] = label;
        }
    };
    if (a.segDone[it.seg]) {
        for (int i = it.start + tid; i < it.end; i += kScanThreads) put(i, -1);
        return;
    }
    const int K = a.segK[it.seg];
    const T rSqr = a.segRSqr[it.seg];
    const T* __restrict__ Sb = a.S + sd.sOff;
    const int64_t ps = a.pStride;
    const int64_t ss = a.sStride;

    T p[PPT][DIM];
    int idx[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        int i = it.start + tid + j * kScanThreads;
        idx[j] = i;
        bool v = i < it.end;
        const int64_t src = v ? point_of(a.pIndex, sd.pOff + i) : 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) p[j][d] = v ? a.P[d * ps + src] : T(0);
    }
    if (K <= 0) {
#pragma unroll
        for (int j = 0; j < PPT; ++j)
            if (idx[j] < it.end) put(idx[j], -1);
        return;
    }

    constexpr T big = FpTraits<T>::big;
    T best[PPT], second[PPT];
    int wt[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        best[j] = big;
        second[j] = big;
        wt[j] = 0;
    }
    const int nT = (K + kChosenTile - 1) / kChosenTile;
    for (int tj = 0; tj < nT; ++tj) {
        T s[kChosenTile][DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
#pragma unroll
            for (int t = 0; t < kChosenTile; ++t) s[t][d] = Sb[d * ss + tj * kChosenTile + t];
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            T m = big;
#pragma unroll
            for (int t = 0; t < kChosenTile; t += 2) {
                T a0 = screen_dist<T, DIM>(p[j], s[t]);
                T a1 = screen_dist<T, DIM>(p[j], s[t + 1]);
                m = tile_min<T>(m, a0, a1);
            }
            const bool c = m < best[j];
            wt[j] = c ? tj : wt[j];
            second[j] = second_update(best[j], second[j], m);
            best[j] = dmin(best[j], m);
        }
    }

    // epilogue: winner inside the winning tile, runner-up, exact decision
    const T delta = T(4 * (DIM + 1)) * FpTraits<T>::unit_roundoff;
    const T tiny = FpTraits<T>::tiny;
    int label[PPT];
    bool amb[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        label[j] = -1;
        amb[j] = false;
        if (idx[j] >= it.end) continue;
        if (!(best[j] <= rSqr * (T(1) + T(2) * delta) + tiny)) continue;   // certainly not a member
        int kwin = -1;
        T secIn = big;
        const int t0 = wt[j] * kChosenTile;
        for (int t = 0; t < kChosenTile; ++t) {
            T sv[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) sv[d] = Sb[d * ss + t0 + t];
            T av = screen_dist<T, DIM>(p[j], sv);
            if (kwin < 0 && av == best[j])
                kwin = t;
            else
                secIn = dmin(secIn, av);
        }
        const T sec = dmin(second[j], secIn);
        if (kwin < 0 || !(sec > best[j] * (T(1) + T(3) * delta) + tiny)) {
            amb[j] = true;
            continue;
        }
        const int k = t0 + kwin;
        T sv[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) sv[d] = Sb[d * ss + k];
        T e = exact_sqdist<T, DIM>(p[j], sv);
        if (e <= rSqr) label[j] = (int)(sd.sOff + k);
    }

    // exact reference scan for near ties (rare)
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        if (__ballot(amb[j]) == 0) continue;
        if (amb[j]) {
            T minSq = big;
            int minK = -1;
            for (int k = 0; k < K; ++k) {
                T sv[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) sv[d] = Sb[d * ss + k];
                T e = exact_sqdist<T, DIM>(p[j], sv);
                if (e < minSq) {
                    minSq = e;
                    minK = k;
                }
            }
            label[j] = (minK >= 0 && minSq <= rSqr) ? (int)(sd.sOff + minK) : -1;
        }
    }
#pragma unroll
    for (int j = 0; j < PPT; ++j)
        if (idx[j] < it.end) put(idx[j], label[j]);
}

// one block per work item
template <typename T, int DIM, int PPT>
__global__ __launch_bounds__(kScanThreads) void evolve_scan_kernel(ScanArgs<T> a) {
    if (*a.allDone) return;
    scan_item<T, DIM, PPT>(a, a.items[blockIdx.x]);
}

// persistent blocks pulling work items from a device queue (items are ordered by
// decreasing cost, so the longest segments start first)
template <typename T, int DIM, int PPT>
__global__ __launch_bounds__(kScanThreads) void evolve_scan_persistent_kernel(ScanArgs<T> a, int numItems) {
    __shared__ int next;
    if (*a.allDone) return;
    if (a.numItemsDev) numItems = *a.numItemsDev;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = next;
        __syncthreads();
        if (item >= numItems) break;
        scan_item<T, DIM, PPT>(a, a.items[item]);
    }
}

// per-iteration reset of the work queue
__global__ void iteration_prologue_kernel(int* __restrict__ queue, const int* __restrict__ allDone) {
    if (*allDone) return;
    *queue = 0;
}

// ---------------------------------------------------------------------------------
// 3. member runs of every chosen slot in the sorted label array
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void chain_bounds_kernel(const int* __restrict__ keys,
                                                           const int64_t* __restrict__ mPtr,
                                                           int* __restrict__ chainOff,
                                                           int* __restrict__ chainEnd,
                                                           const int* __restrict__ allDone) {
    if (*allDone) return;
    const int64_t M = *mPtr;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < M; i += (int64_t)gridDim.x * 256) {
        int k = keys[i];
        if (i == 0 || keys[i - 1] != k) chainOff[k] = (int)i;
        if (i == M - 1 || keys[i + 1] != k) chainEnd[k] = (int)(i + 1);
    }
}

// ---------------------------------------------------------------------------------
// 4a. member coordinates in chain order (contiguous per chosen point)
// ---------------------------------------------------------------------------------
template <typename T, int DIM>
__global__ __launch_bounds__(256) void gather_members_kernel(const T* __restrict__ P, int64_t pStride,
                                                             const int* __restrict__ pIndex,
                                                             const int* __restrict__ members,
                                                             const int64_t* __restrict__ mPtr,
                                                             T* __restrict__ sorted, int64_t mStride,
                                                             const int* __restrict__ allDone) {
    if (*allDone) return;
    const int64_t M = *mPtr;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < M; i += (int64_t)gridDim.x * 256) {
        const int64_t n = point_of(pIndex, members[i]);
#pragma unroll
        for (int d = 0; d < DIM; ++d) sorted[d * mStride + i] = P[d * pStride + n];
    }
}

// Value of lane 16*(lane/16) + J, i.e. lane J of the caller's 16-lane row (DPP
// row_newbcast, gfx90a and later).  Feeding it straight into an add makes the compiler
// emit one v_add_f32_dpp.
template <int J>
__device__ __forceinline__ float row_bcast(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + J, 0xF, 0xF, false));
}
template <int J>
__device__ __forceinline__ double row_bcast(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x150 + J, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x150 + J, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// sum += member J of the current step, for J = 0..15 in order
template <typename T, int J>
__device__ __forceinline__ void fold_row(T& sum, T v) {
    sum = sum + row_bcast<J>(v);
    if constexpr (J < 15) fold_row<T, J + 1>(sum, v);
}

// ---------------------------------------------------------------------------------
// 4b. ordered sums + shift toward the mass centre: one 16-lane row per (coordinate,
// chosen point), coordinate-major, so the coordinates of one point are summed by
// different waves in parallel and a wave serves four points.  Every step the row loads
// 16 consecutive members (lane t: member c0 + t) and folds them into the running sum in
// member order with row broadcasts, i.e. exactly the reference's serial summation
// cordSums[k][d] += P[n][d] in sample order.  Lanes past the end of a chain load +0: the
// running sum starts at +0 and a round-to-nearest sum is -0 only when both addends are
// -0, so it is never -0 and adding +0 leaves it unchanged.  The shift of one coordinate
// depends only on that coordinate's sum (rd/cpu/brute_force/evolve.hpp:129-179).
// ---------------------------------------------------------------------------------
template <typename T, int DIM>
__global__ __launch_bounds__(256) void chain_sum_shift_kernel(ShiftArgs<T> a) {
    if (*a.allDone) return;
    constexpr int kRow = 16;
    const int t = threadIdx.x & (kRow - 1);
    const int64_t row = (int64_t)blockIdx.x * (256 / kRow) + threadIdx.x / kRow;
    const int d = (int)(row / max<int64_t>(a.totalSlots, 1));
    const int64_t slot = row - (int64_t)d * a.totalSlots;
    int seg = -1, off = 0, cnt = 0;
    if (d < DIM) {
        seg = a.slotSeg[slot];
        if (seg >= 0 && !a.segDone[seg] && (int)(slot - a.segs[seg].sOff) < a.segK[seg]) {
            off = a.chainOff[slot];
            cnt = a.chainEnd[slot] - off;
        } else {
            seg = -1;
        }
    }
    // the wave steps through the longest of its four chains; DPP needs every lane present
    int steps = cnt;
    steps = max(steps, __shfl_xor(steps, 16));
    steps = max(steps, __shfl_xor(steps, 32));
    T sum = T(0);
    // Members stream through a ring of kNB buffers of kQ 16-member steps (lane t holds
    // members c + 16 q + t); a buffer is refilled kNB - 1 buffers ahead of its fold, enough
    // to cover the L2 latency with the dependent adds.  Loads are unconditional (index
    // clamped into the chain, or element 0 of the buffer for an empty one) so that no
    // branch separates them from the folds; steps past a chain's end fold +0.
    constexpr int kQ = 1, kNB = 16, kBuf = kQ * kRow;
    const T* __restrict__ src = a.sortedP + (int64_t)min(d, DIM - 1) * a.mStride + off;
    const int last = max(cnt - 1, 0);
    auto load = [&](int i) {
        const T x = src[min(i, last)];
        return i < cnt ? x : T(0);
    };
    const int nBuf = __builtin_amdgcn_readfirstlane((steps + kBuf - 1) / kBuf);   // uniform
    T buf[kNB][kQ];
#pragma unroll
    for (int b = 0; b < kNB; ++b)
#pragma unroll
        for (int q = 0; q < kQ; ++q) buf[b][q] = load(b * kBuf + q * kRow + t);
    for (int g = 0; g * kNB < nBuf; ++g) {
#pragma unroll
        for (int b = 0; b < kNB; ++b) {
#pragma unroll
            for (int q = 0; q < kQ; ++q) fold_row<T, 0>(sum, buf[b][q]);
            const int next = ((g + 1) * kNB + b) * kBuf + t;
#pragma unroll
            for (int q = 0; q < kQ; ++q) buf[b][q] = load(next + q * kRow);
        }
    }
    if (seg < 0 || t != 0) return;
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

// ---------------------------------------------------------------------------------
// 5. convergence bookkeeping (one block)
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void evolve_finalize_kernel(int numSegs, int* __restrict__ segDone,
                                                               int* __restrict__ segMoved,
                                                               int* __restrict__ segIter, int maxIter,
                                                               int* __restrict__ allDone,
                                                               int* __restrict__ totalIters) {
    __shared__ int active;
    if (*allDone) return;
    if (threadIdx.x == 0) active = 0;
    __syncthreads();
    int mine = 0;
    for (int s = threadIdx.x; s < numSegs; s += 1024) {
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

// Writes the far-away sentinel into the slots after K of every segment's last tile.
template <typename T>
__global__ __launch_bounds__(256) void fill_sentinel_kernel(T* __restrict__ S, int64_t sStride, int dim,
                                                            const SegDesc* __restrict__ segs,
                                                            const int* __restrict__ segK, int numSegs) {
    const int seg = blockIdx.x;
    if (seg >= numSegs) return;
    const SegDesc sd = segs[seg];
    const int K = segK[seg];
    const int end = min(sd.sCap, ((K + kChosenTile - 1) / kChosenTile) * kChosenTile);
    for (int k = K + threadIdx.x; k < end; k += 256)
        for (int d = 0; d < dim; ++d) S[d * sStride + sd.sOff + k] = FpTraits<T>::sentinel;
}

// ---------------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------------
template <typename T, int DIM, int PPT>
static void launch_scan_shape(const ScanArgs<T>& a, int numItems, int persistentBlocks, hipStream_t st) {
    if (persistentBlocks > 0) {
        hipLaunchKernelGGL(iteration_prologue_kernel, dim3(1), dim3(1), 0, st, a.queue, a.allDone);
        hipLaunchKernelGGL((evolve_scan_persistent_kernel<T, DIM, PPT>), dim3(std::min(persistentBlocks, numItems)),
                           dim3(kScanThreads), 0, st, a, numItems);
    } else {
        hipLaunchKernelGGL((evolve_scan_kernel<T, DIM, PPT>), dim3(numItems), dim3(kScanThreads), 0, st, a);
    }
}

template <typename T>
hipError_t launch_evolve_scan(const ScanArgs<T>& a, int dim, int numItems, int ppt, int persistentBlocks,
                              hipStream_t st) {
    if (numItems <= 0) return hipSuccess;
    bool ok = false;
    RD_DISPATCH_DIM(dim, {
        if (ppt == 4) {
            launch_scan_shape<T, DIM, 4>(a, numItems, persistentBlocks, st);
            ok = true;
        } else if (ppt == 2) {
            launch_scan_shape<T, DIM, 2>(a, numItems, persistentBlocks, st);
            ok = true;
        } else if (ppt == 1) {
            launch_scan_shape<T, DIM, 1>(a, numItems, persistentBlocks, st);
            ok = true;
        }
    });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_chain_sum_shift(const ShiftArgs<T>& a, int dim, int64_t maxM, bool gather, hipStream_t st) {
    const int64_t blocks = ((int64_t)dim * a.totalSlots + 15) / 16;
    if (blocks <= 0) return hipSuccess;
    const unsigned gblocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((maxM + 255) / 256, 8192));
    bool ok = false;
    RD_DISPATCH_DIM(dim, {
        if (gather)
            hipLaunchKernelGGL((gather_members_kernel<T, DIM>), dim3(gblocks), dim3(256), 0, st, a.P, a.pStride,
                               a.pIndex, a.members, a.mPtr, a.sortedP, a.mStride, a.allDone);
        hipLaunchKernelGGL((chain_sum_shift_kernel<T, DIM>), dim3((unsigned)blocks), dim3(256), 0, st, a);
        ok = true;
    });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

hipError_t launch_chain_bounds(const int* keys, const int64_t* dM, int64_t maxM, int* chainOff,
                               int* chainEnd, const int* allDone, hipStream_t st) {
    int64_t blocks = std::min<int64_t>((maxM + 255) / 256, 8192);
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(chain_bounds_kernel, dim3((unsigned)blocks), dim3(256), 0, st, keys, dM, chainOff,
                       chainEnd, allDone);
    return hipGetLastError();
}

hipError_t launch_evolve_finalize(int numSegs, int* segDone, int* segMoved, int* segIter, int maxIter,
                                  int* allDone, int* totalIters, hipStream_t st) {
    hipLaunchKernelGGL(evolve_finalize_kernel, dim3(1), dim3(1024), 0, st, numSegs, segDone, segMoved,
                       segIter, maxIter, allDone, totalIters);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fill_sentinel(T* S, int64_t sStride, int dim, const SegDesc* segs, const int* segK,
                                int numSegs, hipStream_t st) {
    if (numSegs <= 0) return hipSuccess;
    hipLaunchKernelGGL((fill_sentinel_kernel<T>), dim3(numSegs), dim3(256), 0, st, S, sStride, dim, segs,
                       segK, numSegs);
    return hipGetLastError();
}

template hipError_t launch_evolve_scan<float>(const ScanArgs<float>&, int, int, int, int, hipStream_t);
template hipError_t launch_evolve_scan<double>(const ScanArgs<double>&, int, int, int, int, hipStream_t);
template hipError_t launch_chain_sum_shift<float>(const ShiftArgs<float>&, int, int64_t, bool, hipStream_t);
template hipError_t launch_chain_sum_shift<double>(const ShiftArgs<double>&, int, int64_t, bool, hipStream_t);
template hipError_t launch_fill_sentinel<float>(float*, int64_t, int, const SegDesc*, const int*, int, hipStream_t);
template hipError_t launch_fill_sentinel<double>(double*, int64_t, int, const SegDesc*, const int*, int, hipStream_t);

}  // namespace rdhip
