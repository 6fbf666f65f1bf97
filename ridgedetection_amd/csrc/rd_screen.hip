// rd_screen.hip -- the per-iteration table of the expanded-form evolve screen
// (nearest_labels_x in rd_nearest.hpp): per segment a centre o, the centred chosen points
// s' = s - o and b = |s'|^2, each rounded once, and the bound's B = max b.
#include "rd_evolve.hpp"

namespace rdhip {
namespace {

constexpr int kXThreads = 256;
constexpr int kXWaves = kXThreads / 64;

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// s' = s - o rounded once to T (FP32: the difference of two floats is formed in double, whose
// single rounding to float is the one that counts); b = |s'|^2 rounded once for FP32 (the
// squares of floats are exact in double and their sum rounds at 2^-53), DIM times for FP64.
template <typename T>
__global__ __launch_bounds__(kXThreads) void expand_chosen_kernel(const T* __restrict__ S, int64_t ss,
                                                                 const SegDesc* __restrict__ segs,
                                                                 const int* __restrict__ segK,
                                                                 const int* __restrict__ segDone, int dim,
                                                                 T* __restrict__ X, T* __restrict__ segX,
                                                                 const int* __restrict__ allDone) {
    __shared__ T lo[kMaxDim][kXWaves];
    __shared__ T hi[kMaxDim][kXWaves];
    __shared__ T centre[kMaxDim];
    __shared__ T bw[kXWaves];
    if (*allDone) return;
    const int seg = blockIdx.x;
    if (segDone[seg]) return;
    const SegDesc sd = segs[seg];
    const int K = segK[seg];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const T* __restrict__ Sb = S + sd.sOff;
    T* __restrict__ Xb = X + sd.sOff;
    for (int d = 0; d < dim; ++d) {
        T l = FpTraits<T>::big, h = -FpTraits<T>::big;
        for (int k = tid; k < K; k += kXThreads) {
            const T v = Sb[d * ss + k];
            l = fmin(l, v);
            h = fmax(h, v);
        }
        l = wave_min(l);
        h = wave_max(h);
        if (lane == 0) {
            lo[d][w] = l;
            hi[d][w] = h;
        }
    }
    __syncthreads();
    if (tid < dim) {
        T l = lo[tid][0], h = hi[tid][0];
        for (int i = 1; i < kXWaves; ++i) {
            l = fmin(l, lo[tid][i]);
            h = fmax(h, hi[tid][i]);
        }
        centre[tid] = K > 0 ? T(0.5) * l + T(0.5) * h : T(0);
    }
    __syncthreads();
    const int kEnd = (K + kChosenTile - 1) / kChosenTile * kChosenTile;
    T bmax = T(0);
    for (int k = tid; k < kEnd; k += kXThreads) {
        if (k < K) {
            double acc = 0.0;
            for (int d = 0; d < dim; ++d) {
                const T sp = (T)((double)Sb[d * ss + k] - (double)centre[d]);
                Xb[d * ss + k] = sp;
                acc += (double)sp * (double)sp;
            }
            const T b = (T)acc;
            Xb[dim * ss + k] = b;
            bmax = fmax(bmax, b);
        } else {
            for (int d = 0; d < dim; ++d) Xb[d * ss + k] = T(0);
            Xb[dim * ss + k] = FpTraits<T>::xSentinel;
        }
    }
    bmax = wave_max(bmax);
    if (lane == 0) bw[w] = bmax;
    __syncthreads();
    T* __restrict__ sx = segX + (int64_t)seg * kScreenStride;
    if (tid == 0) {
        T m = bw[0];
        for (int i = 1; i < kXWaves; ++i) m = fmax(m, bw[i]);
        sx[kMaxDim] = m;
    }
    if (tid < dim) sx[tid] = centre[tid];
}

}  // namespace

template <typename T>
hipError_t launch_expand_chosen(const T* S, int64_t sStride, const SegDesc* segs, const int* segK, const int* segDone,
                                int numSegs, int dim, T* xTable, T* segX, const int* allDone, hipStream_t st) {
    if (numSegs <= 0) return hipSuccess;
    if (dim < 1 || dim > kMaxDim) return hipErrorInvalidValue;
    hipLaunchKernelGGL((expand_chosen_kernel<T>), dim3(numSegs), dim3(kXThreads), 0, st, S, sStride, segs, segK,
                       segDone, dim, xTable, segX, allDone);
    return hipGetLastError();
}

template hipError_t launch_expand_chosen<float>(const float*, int64_t, const SegDesc*, const int*, const int*, int, int,
                                                float*, float*, const int*, hipStream_t);
template hipError_t launch_expand_chosen<double>(const double*, int64_t, const SegDesc*, const int*, const int*, int,
                                                 int, double*, double*, const int*, hipStream_t);

}  // namespace rdhip
