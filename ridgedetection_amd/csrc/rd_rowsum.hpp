// rd_rowsum.hpp -- ordered (serial) sums inside 16-lane rows with DPP row broadcasts.
#pragma once

#include "rd_common.hpp"

namespace rdhip {

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

// Longest count over the four 16-lane rows of the wave (wave-uniform).
__device__ __forceinline__ int wave_row_max(int cnt) {
    cnt = max(cnt, __shfl_xor(cnt, 16));
    cnt = max(cnt, __shfl_xor(cnt, 32));
    return __builtin_amdgcn_readfirstlane(cnt);
}

// Serial sum init + m0 + m1 + ... of a row's members in that order: lane t of the row supplies
// member i = 16 q + t through load(i), which must return +0 for i past the row's count
// (a running sum that starts at +0 is never -0, since a round-to-nearest sum is -0 only
// when both addends are -0, so adding +0 leaves it unchanged).  `steps` is the
// wave-uniform longest count.  Members stream through a ring of NB steps refilled NB - 1
// steps ahead of their fold, so the dependent adds cover the load latency.
template <typename T, int NB = 16, typename Load>
__device__ __forceinline__ T row_ordered_sum(int steps, Load&& load, T init = T(0)) {
    constexpr int kRow = 16, kNB = NB;
    const int t = threadIdx.x & (kRow - 1);
    const int nBuf = __builtin_amdgcn_readfirstlane((steps + kRow - 1) / kRow);
    T sum = init;
    T buf[kNB];
#pragma unroll
    for (int b = 0; b < kNB; ++b) buf[b] = load(b * kRow + t);
    for (int g = 0; g * kNB < nBuf; ++g) {
#pragma unroll
        for (int b = 0; b < kNB; ++b) {
            fold_row<T, 0>(sum, buf[b]);
            buf[b] = load(((g + 1) * kNB + b) * kRow + t);
        }
    }
    return sum;
}

}  // namespace rdhip
