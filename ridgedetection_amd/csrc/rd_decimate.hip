// rd_decimate.hip -- order-exact decimation for groups of segments, on gfx950.
//
// Reference semantics:
//   decimate       rd/cpu/brute_force/decimate.hpp:40-81  (brute force, local tiled RD)
//   decimateNoCopy rd/cpu/brute_force/decimate.hpp:84-127 (global decimate of mixed tiled RD)
// Visit the chosen points of a group in order; remove a point when at least 4 live points
// (itself included) lie within r2 or fewer than 3 within 2 r2 (squared 4 r2^2).  Removal
// is immediate, so later points see the smaller set.  decimate stops a pass as soon as
// fewer than 3 points remain; decimateNoCopy instead keeps a point whose removal would
// leave fewer than 3.  Passes repeat until nothing changes or at most 3 points remain.
//
// Device plan (positions do not move during decimation, so the neighbourhoods are fixed):
//   1. layout   -- element e of a group = (segment, k) in segment order (prefix of K);
//   2. count    -- c1(e) = #{j : d2 <= r2^2}, c2(e) = #{j : d2 <= 4 r2^2} over the group
//                  (uniform inner loop: the partner coordinates are scalar loads);
//   3. fill     -- CSR list of the 2 r2 neighbours of every element, with a "within r2" bit;
//   4. scan     -- one wave per group walks the elements in order, 64 at a time: the
//                  window's counters and CSR rows are staged in LDS, decisions are uniform,
//                  and a removal decrements its neighbours' counters (LDS for the window,
//                  device-scope atomics beyond it);
//   5. compact  -- stable in-place compaction of every segment's chosen points.
#include "rd_decimate.hpp"
#include "rd_sort.hpp"

namespace rdhip {

namespace {
constexpr int kRowCap = 8192;   // CSR entries a window stages in LDS

__device__ __forceinline__ int ld_acq(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
}  // namespace

__global__ void dec_copy_k_kernel(const int* __restrict__ segK, int* __restrict__ elemBase, int numSegs) {
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s < numSegs) elemBase[s] = segK[s];
}

__global__ void dec_groups_kernel(DecArgs a) {
    int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= a.numGroups) return;
    const int b = a.groupSegBegin[g], e = a.groupSegEnd[g];
    const int off = a.elemBase[b];
    const int end = e < a.numSegs ? a.elemBase[e] : (int)*a.totalElems;
    a.groupOff[g] = off;
    a.groupCnt[g] = (a.segSkip && a.segSkip[b]) ? 0 : end - off;
}

__global__ __launch_bounds__(256) void dec_elems_kernel(DecArgs a) {
    const int s = blockIdx.x;
    const SegDesc sd = a.segs[s];
    const int K = a.segK[s];
    const int base = a.elemBase[s];
    for (int k = threadIdx.x; k < K; k += 256) {
        a.elemSlot[base + k] = (int)(sd.sOff + k);
        a.elemGroup[base + k] = sd.group;
    }
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void dec_count_kernel(DecArgs a, const T* __restrict__ S, int64_t ss,
                                                        const T* __restrict__ groupRSqr, int fill,
                                                        int64_t maxElems) {
    const int64_t total = *a.totalElems;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) {
        if (!fill && e <= maxElems) a.rowOff[e] = 0;   // tail of the offsets scan
        return;
    }
    const int g = a.elemGroup[e];
    const int o = a.groupOff[g], n = a.groupCnt[g];
    const T rSqr = groupRSqr[g];
    const T r4 = T(4) * rSqr;
    T p[DIM];
    const int se = a.elemSlot[e];
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = S[d * ss + se];
    int c1 = 0, c2 = 0;
    int w = fill ? a.rowOff[e] : 0;
    for (int j = o; j < o + n; ++j) {
        const int sj = a.elemSlot[j];
        T q[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) q[d] = S[d * ss + sj];
        const T dd = exact_sqdist<T, DIM>(p, q);
        const bool in1 = dd <= rSqr;
        const bool in2 = dd <= r4;
        c1 += in1;
        c2 += in2;
        if (fill && in2 && j != e) a.nbr[w++] = (j << 1) | (in1 ? 1 : 0);
    }
    if (!fill) {
        a.c1[e] = c1;
        a.c2[e] = c2;
        // neighbours other than itself (scanned to offsets); elements of a skipped group
        // (n == 0) have no list at all
        a.rowOff[e] = n > 0 ? c2 - 1 : 0;
    }
}

__global__ __launch_bounds__(64) void dec_scan_kernel(DecArgs a) {
    __shared__ int wc1[kWave], wc2[kWave], wal[kWave];
    __shared__ int rows[kRowCap];
    const int g = blockIdx.x;
    const int o = a.groupOff[g], n = a.groupCnt[g];
    const int lane = lane_id();
    for (int e = o + lane; e < o + n; e += kWave) st_rel(&a.alive[e], 1);
    __threadfence();
    int cnt = n;
    int left = -1;
    bool stop = false;
    while (left != cnt && cnt > 3 && !stop) {
        left = cnt;
        for (int w0 = o; w0 < o + n && !stop; w0 += kWave) {
            const int e = w0 + lane;
            const bool inw = e < o + n;
            wc1[lane] = inw ? ld_acq(&a.c1[e]) : 0;
            wc2[lane] = inw ? ld_acq(&a.c2[e]) : 0;
            wal[lane] = inw ? ld_acq(&a.alive[e]) : 0;
            const int wn = min(kWave, o + n - w0);
            const int rs = a.rowOff[w0];
            const int re = a.rowOff[w0 + wn];
            const int myRs = inw ? a.rowOff[e] : 0;
            const int myRe = inw ? a.rowOff[e + 1] : 0;
            const bool lds = (re - rs) <= kRowCap;
            if (lds)
                for (int q = rs + lane; q < re; q += kWave) rows[q - rs] = a.nbr[q];
            __syncthreads();
            for (int q = 0; q < wn; ++q) {
                if (!wal[q]) continue;
                const int x1 = wc1[q], x2 = wc2[q];
                if (!(x1 >= 4 || x2 < 3)) continue;
                if (a.noCopyRule && cnt - 1 < 3) continue;
                --cnt;
                __syncthreads();
                if (lane == 0) wal[q] = 0;
                const int qs = __shfl(myRs, q);
                const int qe = __shfl(myRe, q);
                for (int t = qs + lane; t < qe; t += kWave) {
                    const int v = lds ? rows[t - rs] : a.nbr[t];
                    const int j = v >> 1;
                    const bool close = v & 1;
                    if (j >= w0 && j < w0 + wn) {
                        atomicSub(&wc2[j - w0], 1);
                        if (close) atomicSub(&wc1[j - w0], 1);
                    } else {
                        __hip_atomic_fetch_add(&a.c2[j], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (close)
                            __hip_atomic_fetch_add(&a.c1[j], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                __syncthreads();
                if (!a.noCopyRule && cnt < 3) {
                    stop = true;
                    break;
                }
            }
            if (inw) {
                st_rel(&a.c1[e], wc1[lane]);
                st_rel(&a.c2[e], wc2[lane]);
                st_rel(&a.alive[e], wal[lane]);
            }
            __threadfence();
            __syncthreads();
        }
    }
    if (lane == 0) a.groupLeft[g] = cnt;
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void dec_compact_kernel(DecArgs a, T* __restrict__ S, int64_t ss,
                                                          int* __restrict__ chosenIdx,
                                                          unsigned char* __restrict__ keepOut) {
    __shared__ int waveCnt[4];
    const int s = blockIdx.x;
    if (a.segSkip && a.segSkip[s]) return;
    const SegDesc sd = a.segs[s];
    const int K = a.segK[s];
    const int base = a.elemBase[s];
    const int wave = threadIdx.x / kWave;
    int newK = 0;
    for (int c0 = 0; c0 < K; c0 += 256) {
        const int k = c0 + threadIdx.x;
        const bool live = k < K && ld_acq(&a.alive[base + k]) != 0;
        if (k < K && keepOut) keepOut[sd.sOff + k] = live ? 1 : 0;
        T x[DIM];
        int ci = 0;
        if (live) {
#pragma unroll
            for (int d = 0; d < DIM; ++d) x[d] = S[d * ss + sd.sOff + k];
            if (chosenIdx) ci = chosenIdx[sd.sOff + k];
        }
        const uint64_t m = __ballot(live);
        if (lane_id() == 0) waveCnt[wave] = __popcll(m);
        __syncthreads();
        int before = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            before += w < wave ? waveCnt[w] : 0;
            tot += waveCnt[w];
        }
        if (live) {
            const int dst = newK + before + __popcll(m & lanes_below());
#pragma unroll
            for (int d = 0; d < DIM; ++d) S[d * ss + sd.sOff + dst] = x[d];
            if (chosenIdx) chosenIdx[sd.sOff + dst] = ci;
        }
        newK += tot;
        __syncthreads();
    }
    // sentinel for the rest of the last screening tile
    const int end = min(sd.sCap, ((newK + kChosenTile - 1) / kChosenTile) * kChosenTile);
    for (int k = newK + threadIdx.x; k < end; k += 256)
#pragma unroll
        for (int d = 0; d < DIM; ++d) S[d * ss + sd.sOff + k] = FpTraits<T>::sentinel;
    if (threadIdx.x == 0) a.segK[s] = newK;
}

template <typename T>
hipError_t decimate_groups(DecArgs& a, T* S, int64_t sStride, int dim, const T* groupRSqr, int* chosenIdx,
                           unsigned char* keepOut, int64_t maxElems, DecWorkspace& ws, hipStream_t st) {
    if (a.numSegs <= 0 || a.numGroups <= 0) return hipSuccess;
    hipError_t err;
    // 1. layout
    hipLaunchKernelGGL(dec_copy_k_kernel, dim3((a.numSegs + 255) / 256), dim3(256), 0, st, a.segK, a.elemBase,
                       a.numSegs);
    if ((err = exclusive_scan_inplace(a.elemBase, a.numSegs, ws.partials, a.totalElems, st)) != hipSuccess)
        return err;
    hipLaunchKernelGGL(dec_groups_kernel, dim3((a.numGroups + 255) / 256), dim3(256), 0, st, a);
    hipLaunchKernelGGL(dec_elems_kernel, dim3(a.numSegs), dim3(256), 0, st, a);
    // 2. counts
    const unsigned blocks = (unsigned)((maxElems + 1 + 255) / 256);
    bool ok = false;
    RD_DISPATCH_DIM(dim, {
        hipLaunchKernelGGL((dec_count_kernel<T, DIM>), dim3(blocks), dim3(256), 0, st, a, S, sStride, groupRSqr,
                           0, maxElems);
        ok = true;
    });
    if (!ok) return hipErrorInvalidValue;
    // 3. CSR offsets (rowOff has maxElems + 1 entries; the extra one receives the total)
    if ((err = exclusive_scan_inplace(a.rowOff, maxElems + 1, ws.partials, ws.nbrTotal, st)) != hipSuccess)
        return err;
    int64_t nbrTotal = 0;
    if ((err = hipMemcpyAsync(&nbrTotal, ws.nbrTotal, sizeof(int64_t), hipMemcpyDeviceToHost, st)) != hipSuccess)
        return err;
    if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
    if (nbrTotal > ws.nbrCap) {
        if (ws.nbr) (void)hipFree(ws.nbr);
        ws.nbrCap = nbrTotal + nbrTotal / 4 + 1024;
        if ((err = hipMalloc(&ws.nbr, sizeof(int) * ws.nbrCap)) != hipSuccess) return err;
    }
    a.nbr = ws.nbr;
    RD_DISPATCH_DIM(dim, {
        hipLaunchKernelGGL((dec_count_kernel<T, DIM>), dim3(blocks), dim3(256), 0, st, a, S, sStride, groupRSqr,
                           1, maxElems);
    });
    // 4. ordered removal, one wave per group
    hipLaunchKernelGGL(dec_scan_kernel, dim3(a.numGroups), dim3(kWave), 0, st, a);
    // 5. compaction
    RD_DISPATCH_DIM(dim, {
        hipLaunchKernelGGL((dec_compact_kernel<T, DIM>), dim3(a.numSegs), dim3(256), 0, st, a, S, sStride,
                           chosenIdx, keepOut);
    });
    return hipGetLastError();
}

template hipError_t decimate_groups<float>(DecArgs&, float*, int64_t, int, const float*, int*, unsigned char*,
                                           int64_t, DecWorkspace&, hipStream_t);
template hipError_t decimate_groups<double>(DecArgs&, double*, int64_t, int, const double*, int*, unsigned char*,
                                            int64_t, DecWorkspace&, hipStream_t);

}  // namespace rdhip
