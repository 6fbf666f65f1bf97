// rd_sort.hip -- stable sort of (key, value) int32 pairs by key on gfx950.
//
// LSD radix sort with digits of at most 11 bits.  Each pass is three launches:
//   hist    -- per-block digit histogram (LDS atomics; order-free integer counts)
//   scan    -- exclusive scan of the digit-major histogram [digit][block] (2 kernels)
//   scatter -- per-block stable ranking: every wave ranks its 64 keys with a ballot
//              match (one ballot per digit bit), carries per-digit counters across its
//              chunks in LDS, and the four waves of a block are stitched together by a
//              per-digit prefix, so equal keys keep their input order.  For digits of at
//              most kStageBits bits (runs of at least 4 pairs per digit and block on
//              average) the block then stages its pairs in LDS in block-local sorted order
//              and writes them out in that order: consecutive lanes write consecutive
//              addresses inside each digit's run instead of 64 scattered runs per store.
//              Wider digits leave runs too short to gain: they write straight from the
//              ranks.
// Negative keys are dropped by the first pass, which also writes the count of kept
// pairs.  Used to list the members of every chosen point in point order (evolve) and to
// bin points by tile (tree build).
#include "rd_sort.hpp"

#include <algorithm>

namespace rdhip {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr int kChunks = 8;                                   // 64-key chunks per wave
constexpr int kBlockElems = kWavesPerBlock * kChunks * kWave;  // 2048
constexpr int kMaxDigitBits = 11;
constexpr int kScanElems = 4096;                              // entries per scan block

__global__ __launch_bounds__(kThreads) void radix_hist_kernel(const int* __restrict__ keys,
                                                              const int64_t* __restrict__ nPtr,
                                                              int shift, int bits, int numBlocks,
                                                              int* __restrict__ ghist,
                                                              const int* __restrict__ skip) {
    __shared__ int h[1 << kMaxDigitBits];
    if (skip && *skip) return;
    const int nb = 1 << bits;
    for (int v = threadIdx.x; v < nb; v += kThreads) h[v] = 0;
    __syncthreads();
    const int64_t n = *nPtr;
    const int64_t b0 = (int64_t)blockIdx.x * kBlockElems;
    const int mask = nb - 1;
    for (int e = threadIdx.x; e < kBlockElems; e += kThreads) {
        int64_t i = b0 + e;
        if (i < n) {
            int k = keys[i];
            if (k >= 0) atomicAdd(&h[(k >> shift) & mask], 1);
        }
    }
    __syncthreads();
    for (int v = threadIdx.x; v < nb; v += kThreads) ghist[(int64_t)v * numBlocks + blockIdx.x] = h[v];
}

// Exclusive scan, phase 1: per-block totals of kScanElems entries.
__global__ __launch_bounds__(kThreads) void scan_reduce_kernel(const int* __restrict__ a, int64_t L,
                                                               int* __restrict__ partials,
                                                               const int* __restrict__ skip) {
    __shared__ int red[kThreads];
    if (skip && *skip) return;
    const int64_t b0 = (int64_t)blockIdx.x * kScanElems;
    int s = 0;
    for (int e = threadIdx.x; e < kScanElems; e += kThreads) {
        int64_t i = b0 + e;
        if (i < L) s += a[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// Phase 2: scan inside each block (16 consecutive entries per thread) plus its offset, the
// sum of the partials of the blocks before it (summed here: a separate scan of the partials
// would cost a launch); the last block writes the grand total.
__global__ __launch_bounds__(kThreads) void scan_apply_kernel(int* __restrict__ a, int64_t L,
                                                              const int* __restrict__ partials,
                                                              int64_t* __restrict__ total,
                                                              const int* __restrict__ skip) {
    __shared__ int buf[kThreads];
    if (skip && *skip) return;
    int c = 0;
    for (int q = threadIdx.x; q < (int)blockIdx.x; q += kThreads) c += partials[q];
    buf[threadIdx.x] = c;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) buf[threadIdx.x] += buf[threadIdx.x + o];
        __syncthreads();
    }
    const int carry = buf[0];
    __syncthreads();
    constexpr int per = kScanElems / kThreads;
    const int64_t b0 = (int64_t)blockIdx.x * kScanElems + (int64_t)threadIdx.x * per;
    int v[per];
    int s = 0;
#pragma unroll
    for (int j = 0; j < per; ++j) {
        int64_t i = b0 + j;
        v[j] = i < L ? a[i] : 0;
        s += v[j];
    }
    buf[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1) {
        int t = threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
        __syncthreads();
        buf[threadIdx.x] += t;
        __syncthreads();
    }
    int run = carry + buf[threadIdx.x] - s;
#pragma unroll
    for (int j = 0; j < per; ++j) {
        int64_t i = b0 + j;
        if (i < L) a[i] = run;
        run += v[j];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kThreads - 1) *total = run;
}

// Where a ranked pair goes: its value into valsOut ...
struct ValueSink {
    int* vals;
    __device__ __forceinline__ void put(int64_t dst, int v) const { vals[dst] = v; }
};
// ... or the coordinates of the point it names into a SoA output (last pass of the evolve sort)
template <typename T, int DIM>
struct GatherSink {
    const T* P;
    int64_t pStride;
    const int* pIndex;
    T* out;
    int64_t outStride;
    __device__ __forceinline__ void put(int64_t dst, int v) const {
        const int64_t n = point_of(pIndex, v);
#pragma unroll
        for (int d = 0; d < DIM; ++d) out[d * outStride + dst] = P[d * pStride + n];
    }
};

constexpr int kStageBits = 9;

template <bool kStage, typename Sink>
__global__ __launch_bounds__(kThreads) void radix_scatter_kernel(
    const int* __restrict__ keysIn, const int* __restrict__ valsIn, const int64_t* __restrict__ nPtr,
    int shift, int bits, int numBlocks, const int* __restrict__ ghistScanned,
    int* __restrict__ keysOut, Sink sink, const int* __restrict__ skip) {
    constexpr int kDigits = 1 << (kStage ? kStageBits : kMaxDigitBits);
    constexpr int kPer = (kDigits + kThreads - 1) / kThreads;
    // per-wave digit counters, later reused for the staged (destination, key, value) triples
    constexpr int kSmem = kStage ? std::max(kWavesPerBlock * kDigits, 3 * kBlockElems) : kWavesPerBlock * kDigits;
    __shared__ int smem[kSmem];
    __shared__ int lbase[kStage ? kDigits : 1];
    __shared__ int wsum[kWavesPerBlock];
    int (*cnt)[kDigits] = reinterpret_cast<int (*)[kDigits]>(smem);
    if (skip && *skip) return;
    const int nb = 1 << bits;
    const int mask = nb - 1;
    for (int v = threadIdx.x; v < nb; v += kThreads)
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) cnt[w][v] = 0;
    __syncthreads();

    const int64_t n = *nPtr;
    const int wave = threadIdx.x / kWave;
    const int lane = lane_id();
    const uint64_t below = lanes_below();
    const int64_t w0 = (int64_t)blockIdx.x * kBlockElems + (int64_t)wave * (kChunks * kWave);

    int key[kChunks], pos[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
        int64_t i = w0 + c * kWave + lane;
        int k = (i < n) ? keysIn[i] : -1;
        key[c] = k;
        const bool valid = k >= 0;
        const int dig = valid ? ((k >> shift) & mask) : 0;
        uint64_t m = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (dig >> b) & 1;
            const uint64_t bb = __ballot(valid && bit);
            m &= bit ? bb : ~bb;
        }
        int p = 0;
        if (valid) {
            p = cnt[wave][dig] + __popcll(m & below);
        }
        // the lowest lane of each group of equal digits advances that digit's counter;
        // within a wave LDS operations complete in issue order, so the reads above
        // observe the counter before this update.
        if (valid && (m & below) == 0) cnt[wave][dig] += __popcll(m);
        pos[c] = p;
    }
    __syncthreads();
    if constexpr (!kStage) {
        for (int v = threadIdx.x; v < nb; v += kThreads) {
            int run = ghistScanned[(int64_t)v * numBlocks + blockIdx.x];
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; ++w) {
                int t = cnt[w][v];
                cnt[w][v] = run;
                run += t;
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const int k = key[c];
            if (k >= 0) {
                const int64_t i = w0 + c * kWave + lane;
                const int d = cnt[wave][(k >> shift) & mask] + pos[c];
                keysOut[d] = k;
                sink.put(d, valsIn ? valsIn[i] : (int)i);
            }
        }
    } else {
        // block-local digit bases: exclusive scan of the block's per-digit totals
        int tot[kPer];
        int tsum = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int v = threadIdx.x * kPer + j;
            int t = 0;
            if (v < nb)
#pragma unroll
                for (int w = 0; w < kWavesPerBlock; ++w) t += cnt[w][v];
            tot[j] = t;
            tsum += t;
        }
        int incl = tsum;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == kWave - 1) wsum[wave] = incl;
        __syncthreads();
        int excl = incl - tsum;
        int blockCount = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) {
            excl += w < wave ? wsum[w] : 0;
            blockCount += wsum[w];
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int v = threadIdx.x * kPer + j;
            if (v < nb) lbase[v] = excl;
            excl += tot[j];
        }
        // global bases per (wave, digit)
        for (int v = threadIdx.x; v < nb; v += kThreads) {
            int run = ghistScanned[(int64_t)v * numBlocks + blockIdx.x];
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; ++w) {
                int t = cnt[w][v];
                cnt[w][v] = run;
                run += t;
            }
        }
        __syncthreads();
        int dst[kChunks], rank[kChunks];
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const int k = key[c];
            dst[c] = 0;
            rank[c] = 0;
            if (k >= 0) {
                const int dig = (k >> shift) & mask;
                dst[c] = cnt[wave][dig] + pos[c];
                rank[c] = lbase[dig] + (cnt[wave][dig] - cnt[0][dig]) + pos[c];
            }
        }
        __syncthreads();   // the counters are dead: stage the triples in block-local order
        int* sDst = smem;
        int* sKey = smem + kBlockElems;
        int* sVal = smem + 2 * kBlockElems;
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const int k = key[c];
            if (k >= 0) {
                const int64_t i = w0 + c * kWave + lane;
                sDst[rank[c]] = dst[c];
                sKey[rank[c]] = k;
                sVal[rank[c]] = valsIn ? valsIn[i] : (int)i;
            }
        }
        __syncthreads();
        for (int r = threadIdx.x; r < blockCount; r += kThreads) {
            const int d = sDst[r];
            keysOut[d] = sKey[r];
            sink.put(d, sVal[r]);
        }
    }
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

template <typename Sink>
void launch_scatter(const int* inK, const int* inV, const int64_t* n, int shift, int w, int blocks, const int* ghist,
                    int* oK, const Sink& sink, const int* skip, hipStream_t stream) {
    if (w <= kStageBits)
        hipLaunchKernelGGL((radix_scatter_kernel<true, Sink>), dim3((unsigned)blocks), dim3(kThreads), 0, stream, inK,
                           inV, n, shift, w, blocks, ghist, oK, sink, skip);
    else
        hipLaunchKernelGGL((radix_scatter_kernel<false, Sink>), dim3((unsigned)blocks), dim3(kThreads), 0, stream, inK,
                           inV, n, shift, w, blocks, ghist, oK, sink, skip);
}

}  // namespace

int64_t scan_partials_count(int64_t L) { return cdiv(std::max<int64_t>(L, 1), kScanElems); }

hipError_t exclusive_scan_inplace(int* a, int64_t L, int* partials, int64_t* total, hipStream_t stream) {
    if (L <= 0) return hipSuccess;
    const int64_t scanBlocks = cdiv(L, kScanElems);
    hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)scanBlocks), dim3(kThreads), 0, stream, a, L, partials,
                       nullptr);
    hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)scanBlocks), dim3(kThreads), 0, stream, a, L, partials,
                       total, nullptr);
    return hipGetLastError();
}

bool sort_single_pass(int numKeys) { return numKeys <= (1 << kMaxDigitBits); }

namespace {
__global__ __launch_bounds__(256) void key_ranges_kernel(const int* __restrict__ ghist, int numBlocks,
                                                         int numDigits, int numKeys, const int64_t* __restrict__ dM,
                                                         int* __restrict__ off, int* __restrict__ end,
                                                         int* __restrict__ longList, int longShift,
                                                         const int* __restrict__ skip) {
    if (skip && *skip) return;
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= numKeys) return;
    const int64_t M = *dM;
    const int64_t o = ghist[(int64_t)k * numBlocks];
    const int64_t e = k + 1 < numDigits ? (int64_t)ghist[(int64_t)(k + 1) * numBlocks] : M;
    off[k] = (int)o;
    end[k] = (int)e;
    if (longList && e > o) {
        const int64_t L = int64_t(1) << longShift;
        const int64_t m0 = (o + L - 1) >> longShift;
        for (int64_t m = m0; (m << longShift) < e; ++m) longList[m] = m == m0 ? k : -1;
    }
}
}  // namespace

hipError_t launch_key_ranges_from_sort(const void* workspace, int64_t maxN, int numKeys, const int64_t* dM,
                                       int* off, int* end, int* longList, int longShift, const int* skip,
                                       hipStream_t st) {
    if (!sort_single_pass(numKeys)) return hipErrorInvalidValue;
    int bits = 1;
    while ((int64_t(1) << bits) < numKeys) ++bits;
    const int64_t blocks = cdiv(std::max<int64_t>(maxN, 1), kBlockElems);
    hipLaunchKernelGGL(key_ranges_kernel, dim3((unsigned)((numKeys + 255) / 256)), dim3(256), 0, st,
                       static_cast<const int*>(workspace), (int)blocks, 1 << bits, numKeys, dM, off, end, longList,
                       longShift, skip);
    return hipGetLastError();
}

size_t sort_workspace_bytes(int64_t maxN, int numKeys) {
    int bits = 1;
    while ((int64_t(1) << bits) < numKeys) ++bits;
    int passes = (bits + kMaxDigitBits - 1) / kMaxDigitBits;
    int w = (bits + passes - 1) / passes;
    int64_t blocks = cdiv(std::max<int64_t>(maxN, 1), kBlockElems);
    int64_t L = (int64_t(1) << w) * blocks;
    int64_t scanBlocks = cdiv(L, kScanElems);
    // ghist + partials + temp keys/vals + count
    return sizeof(int) * (size_t)(L + scanBlocks + 2 * std::max<int64_t>(maxN, 1)) + 64;
}

namespace {

// LSD passes; `lastScatter(inK, inV, n, shift, w, blocks, ghist)` launches the final pass.
template <typename Last>
hipError_t sort_passes(const int* keysIn, const int* valsIn, const int64_t* dN, int64_t maxN, int numKeys,
                       int* keysOut, int* valsOut, int64_t* dM, void* workspace, hipStream_t stream,
                       const int* skip, Last&& lastScatter) {
    if (maxN <= 0) return hipSuccess;
    int bits = 1;
    while ((int64_t(1) << bits) < numKeys) ++bits;
    const int passes = (bits + kMaxDigitBits - 1) / kMaxDigitBits;
    const int w = (bits + passes - 1) / passes;
    const int64_t blocks = cdiv(maxN, kBlockElems);
    const int64_t L = (int64_t(1) << w) * blocks;
    const int64_t scanBlocks = cdiv(L, kScanElems);
    int* ghist = static_cast<int*>(workspace);
    int* partials = ghist + L;
    int* tmpK = partials + scanBlocks;
    int* tmpV = tmpK + maxN;

    const int* inK = keysIn;
    const int* inV = valsIn;
    const int64_t* n = dN;
    for (int p = 0; p < passes; ++p) {
        // ping-pong so that the last pass lands in keysOut/valsOut
        const bool toOut = ((passes - 1 - p) % 2) == 0;
        int* oK = toOut ? keysOut : tmpK;
        int* oV = toOut ? valsOut : tmpV;
        const int shift = p * w;
        hipLaunchKernelGGL(radix_hist_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, stream, inK, n,
                           shift, w, (int)blocks, ghist, skip);
        hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)scanBlocks), dim3(kThreads), 0, stream,
                           ghist, L, partials, skip);
        hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)scanBlocks), dim3(kThreads), 0, stream,
                           ghist, L, partials, p == 0 ? dM : nullptr, skip);
        if (p == passes - 1)
            lastScatter(inK, inV, n, shift, w, (int)blocks, ghist, oK, oV);
        else
            launch_scatter(inK, inV, n, shift, w, (int)blocks, ghist, oK, ValueSink{oV}, skip, stream);
        inK = oK;
        inV = oV;
        n = dM;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t stable_sort_by_key(const int* keysIn, const int* valsIn, const int64_t* dN, int64_t maxN,
                              int numKeys, int* keysOut, int* valsOut, int64_t* dM, void* workspace,
                              hipStream_t stream, const int* skip) {
    return sort_passes(keysIn, valsIn, dN, maxN, numKeys, keysOut, valsOut, dM, workspace, stream, skip,
                       [&](const int* inK, const int* inV, const int64_t* n, int shift, int w, int blocks,
                           const int* ghist, int* oK, int* oV) {
                           launch_scatter(inK, inV, n, shift, w, blocks, ghist, oK, ValueSink{oV}, skip, stream);
                       });
}

template <typename T>
hipError_t stable_sort_by_key(const int* keysIn, const int* valsIn, const int64_t* dN, int64_t maxN,
                              int numKeys, int* keysOut, int* valsOut, int64_t* dM, void* workspace,
                              hipStream_t stream, const int* skip, const SortGather<T>& g) {
    bool ok = false;
    hipError_t e = sort_passes(
        keysIn, valsIn, dN, maxN, numKeys, keysOut, valsOut, dM, workspace, stream, skip,
        [&](const int* inK, const int* inV, const int64_t* n, int shift, int w, int blocks, const int* ghist,
            int* oK, int*) {
            RD_DISPATCH_DIM(g.dim, {
                launch_scatter(inK, inV, n, shift, w, blocks, ghist, oK,
                               GatherSink<T, DIM>{g.P, g.pStride, g.pIndex, g.out, g.outStride}, skip, stream);
                ok = true;
            });
        });
    if (e != hipSuccess) return e;
    return ok || maxN <= 0 ? hipSuccess : hipErrorInvalidValue;
}

template hipError_t stable_sort_by_key<float>(const int*, const int*, const int64_t*, int64_t, int, int*, int*,
                                              int64_t*, void*, hipStream_t, const int*, const SortGather<float>&);
template hipError_t stable_sort_by_key<double>(const int*, const int*, const int64_t*, int64_t, int, int*, int*,
                                               int64_t*, void*, hipStream_t, const int*, const SortGather<double>&);

}  // namespace rdhip
