// rd_scan_impl.hpp -- the evolve scan kernels (nearest chosen point of every sample) and
// their launcher, instantiated per value type by rd_scan_f32.hip / rd_scan_f64.hip so the
// (dimension x samples per lane x block size) variants compile in parallel.
#pragma once

#include "rd_chain.hpp"
#include "rd_nearest.hpp"

namespace rdhip {

// ---------------------------------------------------------------------------------
// 1. nearest chosen point
// ---------------------------------------------------------------------------------
// lh (optional): the block's histogram of the labels below lhKeys, in LDS.  F: the screen form
// (rd_nearest.hpp): 0 direct, 1 expanded (a.xTable), 2 expanded on the matrix cores (FP32;
// wave w of the block takes the item's samples [64 PPT w, 64 PPT (w + 1)) in column blocks of 32)
template <typename T, int DIM, int PPT, int TH, int F>
__device__ __forceinline__ void scan_item(const ScanArgs<T>& a, const WorkItem it, int* lh = nullptr,
                                          int lhKeys = 0) {
    const SegDesc sd = a.segs[it.seg];
    const int tid = threadIdx.x;
    int* __restrict__ lab = a.labels + sd.pOff;
    // compact output: slot of sample i is it.pad + (i - it.start)
    int* __restrict__ ck = a.outKeys ? a.outKeys + it.pad - it.start : nullptr;
    int* __restrict__ cp = a.outKeys ? a.outPos + it.pad - it.start : nullptr;
    auto put = [&](int i, int label) {
        if (ck) {
            ck[i] = (label >= 0 && a.denseOf) ? a.denseOf[label] : label;
            cp[i] = (int)(sd.pOff + i);
        } else {
            lab[i] = label;
        }
    };
    if (a.segDone[it.seg]) {
        for (int i = it.start + tid; i < it.end; i += TH) put(i, -1);
        return;
    }
    const int K = a.segK[it.seg];
    if constexpr (F == 2) {
        static_assert(sizeof(T) == 4, "matrix-core screen is FP32");
        constexpr int CB = 2 * PPT;
        const int lane = tid & 63, h = lane >> 5;
        const int base = it.start + (tid >> 6) * (CB * 32);
        T pc[CB][DIM];
        bool vc[CB];
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            const int i = base + c * 32 + (lane & 31);
            vc[c] = i < it.end;
            const int64_t src = vc[c] ? point_of(a.pIndex, sd.pOff + i) : 0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) pc[c][d] = vc[c] ? a.P[d * a.pStride + src] : T(0);
        }
        const T* __restrict__ sx = a.segX + (int64_t)it.seg * kScreenStride;
        T o[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) o[d] = sx[d];
        int label[PPT];
        nearest_labels_mfma<DIM, CB>(pc, vc, K, a.S + sd.sOff, a.xTable + sd.sOff, a.sStride, o, sx[kMaxDim],
                                     a.segRSqr[it.seg], label);
#pragma unroll
        for (int e = 0; e < PPT; ++e) {
            const int i = base + (2 * e + h) * 32 + (lane & 31);
            if (i < it.end) {
                put(i, label[e] < 0 ? -1 : (int)(sd.sOff + label[e]));
                if (lh && label[e] >= 0 && label[e] < lhKeys) atomicAdd(&lh[label[e]], 1);
            }
        }
        return;
    }
    T p[PPT][DIM];
    bool valid[PPT];
    int idx[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        int i = it.start + tid + j * TH;
        idx[j] = i;
        valid[j] = i < it.end;
        const int64_t src = valid[j] ? point_of(a.pIndex, sd.pOff + i) : 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) p[j][d] = valid[j] ? a.P[d * a.pStride + src] : T(0);
    }
    int label[PPT];
    if constexpr (F == 1) {
        const T* __restrict__ sx = a.segX + (int64_t)it.seg * kScreenStride;
        T o[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) o[d] = sx[d];
        nearest_labels_x<T, DIM, PPT>(p, valid, K, a.S + sd.sOff, a.xTable + sd.sOff, a.sStride, o, sx[kMaxDim],
                                      a.segRSqr[it.seg], label);
    } else {
        nearest_labels<T, DIM, PPT>(p, valid, K, a.S + sd.sOff, a.sStride, a.segRSqr[it.seg], label);
    }
#pragma unroll
    for (int j = 0; j < PPT; ++j)
        if (valid[j]) {
            put(idx[j], label[j] < 0 ? -1 : (int)(sd.sOff + label[j]));
            if (lh && label[j] >= 0 && label[j] < lhKeys) atomicAdd(&lh[label[j]], 1);
        }
}

// one block per work item (blocks past a device-side item count exit)
template <typename T, int DIM, int PPT, int TH, int F>
__global__ __launch_bounds__(TH) void evolve_scan_kernel(ScanArgs<T> a) {
    if (*a.allDone) return;
    if (a.numItemsDev && (int)blockIdx.x >= *a.numItemsDev) return;
    scan_item<T, DIM, PPT, TH, F>(a, a.items[blockIdx.x]);
}

// one block per item in use, in XCD-aware order, and the histogram of the item's labels
// (slot - sOff below the segment's key capacity) into its segment's count table (rd_chain.hpp)
template <typename T, int DIM, int PPT, int TH, int F>
__global__ __launch_bounds__(TH) void evolve_scan_hist_kernel(ScanArgs<T> a) {
    __shared__ int lh[kChainMaxKeys];
    if (*a.allDone) return;
    const int n = a.numItemsDev ? *a.numItemsDev : (int)gridDim.x;
    if ((int)blockIdx.x >= n) return;
    const WorkItem it = a.items[xcd_item(blockIdx.x, n)];
    const int K = min(min(a.segK[it.seg], a.histKeyCap[it.seg]), kChainMaxKeys);
    for (int k = threadIdx.x; k < K; k += TH) lh[k] = 0;
    __syncthreads();
    scan_item<T, DIM, PPT, TH, F>(a, it, lh, K);
    __syncthreads();
    const int nb = (a.segs[it.seg].nP + a.histItem - 1) / a.histItem;
    int* __restrict__ col = a.blockHist + a.histSegOff[it.seg] + it.start / a.histItem;
    for (int k = threadIdx.x; k < K; k += TH) col[(int64_t)k * nb] = lh[k];
}

// persistent blocks pulling work items from a device queue (items are ordered by
// decreasing cost, so the longest segments start first)
template <typename T, int DIM, int PPT, int F>
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
        scan_item<T, DIM, PPT, kScanThreads, F>(a, a.items[item]);
    }
}

// per-iteration reset of the work queue of the persistent form
namespace {
__global__ void iteration_prologue_kernel(int* __restrict__ queue, const int* __restrict__ allDone) {
    if (*allDone) return;
    *queue = 0;
}
}  // namespace

template <typename T, int DIM, int PPT, int TH, int F>
void launch_scan_shape(const ScanArgs<T>& a, int numItems, int persistentBlocks, hipStream_t st) {
    if (a.blockHist) {
        if constexpr (DIM <= kChainMaxDim)
            hipLaunchKernelGGL((evolve_scan_hist_kernel<T, DIM, PPT, TH, F>), dim3(numItems), dim3(TH), 0, st, a);
    } else if (persistentBlocks > 0) {
        hipLaunchKernelGGL(iteration_prologue_kernel, dim3(1), dim3(1), 0, st, a.queue, a.allDone);
        hipLaunchKernelGGL((evolve_scan_persistent_kernel<T, DIM, PPT, F>), dim3(std::min(persistentBlocks, numItems)),
                           dim3(kScanThreads), 0, st, a, numItems);
    } else {
        hipLaunchKernelGGL((evolve_scan_kernel<T, DIM, PPT, TH, F>), dim3(numItems), dim3(TH), 0, st, a);
    }
}

template <typename T, int DIM, int PPT, int F>
bool launch_scan_threads(const ScanArgs<T>& a, int numItems, int threads, int persistentBlocks, hipStream_t st) {
    if (persistentBlocks > 0 || threads == 256) {
        launch_scan_shape<T, DIM, PPT, 256, F>(a, numItems, persistentBlocks, st);
    } else if (threads == 128) {
        launch_scan_shape<T, DIM, PPT, 128, F>(a, numItems, 0, st);
    } else if (threads == 512) {
        launch_scan_shape<T, DIM, PPT, 512, F>(a, numItems, 0, st);
    } else {
        return false;
    }
    return true;
}

// F: the screen form (scan_item; a.xTable and a.segX set for F > 0); each value type
// instantiates every form in a translation unit of its own (rd_scan_f32.hip,
// rd_scan_x_f32.hip, rd_scan_m_f32.hip, ...)
template <typename T, int F>
hipError_t launch_evolve_scan_impl(const ScanArgs<T>& a, int dim, int numItems, int ppt, int threads,
                                   int persistentBlocks, hipStream_t st) {
    if (numItems <= 0) return hipSuccess;
    if (a.blockHist && (dim > kChainMaxDim || persistentBlocks > 0)) return hipErrorInvalidValue;
    if ((F > 0) != (a.xTable != nullptr) || (F > 0 && !a.segX) || (F == 2) != (a.mfma != 0))
        return hipErrorInvalidValue;
    bool ok = false;
    RD_DISPATCH_DIM(dim, {
        if (ppt == 4)
            ok = launch_scan_threads<T, DIM, 4, F>(a, numItems, threads, persistentBlocks, st);
        else if (ppt == 2)
            ok = launch_scan_threads<T, DIM, 2, F>(a, numItems, threads, persistentBlocks, st);
        else if (ppt == 1)
            ok = launch_scan_threads<T, DIM, 1, F>(a, numItems, threads, persistentBlocks, st);
    });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

// the expanded-form instantiations (rd_scan_x_f32.hip, rd_scan_x_f64.hip) and the matrix-core
// ones (rd_scan_m_f32.hip; FP32 only)
template <typename T>
hipError_t launch_evolve_scan_x(const ScanArgs<T>& a, int dim, int numItems, int ppt, int threads,
                                int persistentBlocks, hipStream_t st);
hipError_t launch_evolve_scan_m(const ScanArgs<float>& a, int dim, int numItems, int ppt, int threads,
                                int persistentBlocks, hipStream_t st);

}  // namespace rdhip
