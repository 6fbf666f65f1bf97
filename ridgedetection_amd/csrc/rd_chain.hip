// rd_chain.hip -- member ordering for the evolve iteration of one segment with at most
// kChainMaxKeys chosen points (a whole cloud), after the scan (rd_scan_impl.hpp) has written
// every sample's label and every block's label histogram.
//
// Reference semantics: rd/cpu/brute_force/evolve.hpp:32-77 (the members of every chosen point,
// summed in sample order).
//
//   key_prefix     -- one wave per key: the key's row of block counts becomes its exclusive
//                     prefix over blocks, and the key's member total.
//   member_scatter -- one block per scan item (same samples, same XCD-aware order): ranks the
//                     item's samples by key in sample order (ballot match per key bit, as in
//                     rd_sort.hip), reads their coordinates coalesced, and writes them in chain
//                     order through an LDS stage, so consecutive threads write consecutive
//                     addresses of each chain run.  Chain k lies in [chainOff[k], chainEnd[k])
//                     with chainEnd[k] the running sum of the totals rounded up to 16 bytes.
//                     Block 0 also writes the chain ranges and the wave-summed chain list of the
//                     ordered sums (chain_sum_shift_kernel, rd_evolve.hip), which follow.
// This replaces the histogram pass, the two-kernel histogram scan, the sort's scatter and the
// chain-range pass of the general path (rd_sort.hip + chain bounds) for this case.
#include "rd_chain.hpp"

#include <algorithm>

namespace rdhip {

namespace {

// Exclusive scan over the block of one value per thread (thread order); *total = the sum.
template <int W>
__device__ __forceinline__ int block_excl_scan(int v, int* wsum, int& total) {
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    int incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == kWave - 1) wsum[wave] = incl;
    __syncthreads();
    int excl = incl - v;
    total = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
        excl += w < wave ? wsum[w] : 0;
        total += wsum[w];
    }
    __syncthreads();
    return excl;
}

// ---------------------------------------------------------------------------------
// per-key prefix over blocks
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void key_prefix_kernel(int* __restrict__ hist, int64_t histStride, int nb,
                                                         const int* __restrict__ segK, int kb,
                                                         int* __restrict__ totals, const int* __restrict__ allDone) {
    if (*allDone) return;
    const int lane = lane_id();
    const int k = blockIdx.x * 4 + threadIdx.x / kWave;
    if (k >= kb) return;
    int* __restrict__ row = hist + (int64_t)k * histStride;
    constexpr int kPer = 16;   // one pass for up to 1024 blocks
    int carry = 0;
    for (int base = 0; base < nb; base += kWave * kPer) {
        const int i0 = base + lane * kPer;
        int v[kPer];
        int s = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            v[j] = i0 + j < nb ? row[i0 + j] : 0;
            s += v[j];
        }
        // rows of keys past the live count hold stale counts: loaded (in bounds, in parallel
        // with the count's load) but neither rewritten nor totalled
        if (k >= min(segK[0], kb)) {
            if (lane == 0) totals[k] = 0;
            return;
        }
        int incl = s;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        int run = carry + incl - s;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            if (i0 + j < nb) row[i0 + j] = run;
            run += v[j];
        }
        carry += __shfl(incl, kWave - 1);
    }
    if (lane == 0) totals[k] = carry;
}

// ---------------------------------------------------------------------------------
// member scatter: W waves, C chunks of 64 samples per wave (item size W * C * 64)
// ---------------------------------------------------------------------------------
constexpr int kMaxKeysPerThread = kChainMaxKeys / 128;

template <typename T, int DIM, int W, int C>
__global__ __launch_bounds__(W * kWave) void member_scatter_kernel(ChainArgs<T> a, int keyCap) {
    constexpr int TH = W * kWave;
    constexpr int E = TH * C;
    constexpr int A = kChainAlign<T>;
    static_assert(kChainMaxKeys / TH <= kMaxKeysPerThread, "keys per thread");
    extern __shared__ __align__(16) int lds[];
    // [0, region): per-wave key counters cnt[w * keyCap + k], later the staged destinations
    // and values; then gb[keyCap] (global base of this block's run of every key),
    // lbase[keyCap] (block-local base of every key) and wsum[W]
    constexpr int kStage = E + E * (int)(sizeof(T) / sizeof(int));
    const int region = W * keyCap > kStage ? W * keyCap : kStage;
    int* cnt = lds;
    int* gb = lds + region;
    int* lbase = gb + keyCap;
    int* wsum = lbase + keyCap;
    if (*a.allDone) return;
    const int K = min(a.segK[0], a.keyBound);
    const int b = xcd_item(blockIdx.x, gridDim.x);
    const WorkItem it = a.items[b];
    const SegDesc sd = a.segs[it.seg];
    const int tid = threadIdx.x, wave = tid / kWave, lane = lane_id();
    const int kpt = keyCap / TH;

    // every global load of the block first, so that their latencies overlap: the item's labels
    // and coordinates (wave w holds chunks w*C .. w*C+C-1), the key totals and this block's
    // column of the per-key block prefixes
    const int* __restrict__ lab = a.labels + sd.pOff;
    int lraw[C];
    T p[C][DIM];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int i = it.start + (wave * C + c) * kWave + lane;
        const bool in = i < it.end;
        lraw[c] = in ? lab[i] : -1;
        const int64_t n = point_of(a.pIndex, sd.pOff + (in ? i : it.start));
#pragma unroll
        for (int d = 0; d < DIM; ++d) p[c][d] = a.P[d * a.pStride + n];
    }
    int tk[kMaxKeysPerThread], hv[kMaxKeysPerThread];
#pragma unroll
    for (int j = 0; j < kMaxKeysPerThread; ++j) {
        const int k = tid * kpt + j;
        const bool live = j < kpt && k < K;
        tk[j] = live ? a.totals[k] : 0;
        hv[j] = live ? a.hist[(int64_t)k * a.histStride + b] : 0;
    }

    for (int v = tid; v < W * keyCap; v += TH) cnt[v] = 0;
    if (blockIdx.x == 0)
        for (int e = tid; e < a.longEntries; e += TH) a.longList[e] = -1;
    // chain k ends at the running sum of the totals rounded up to A values
    {
        int sk[kMaxKeysPerThread];
        int sum = 0;
#pragma unroll
        for (int j = 0; j < kMaxKeysPerThread; ++j) {
            sk[j] = (tk[j] + A - 1) / A * A;
            sum += sk[j];
        }
        int all;
        int end = block_excl_scan<W>(sum, wsum, all);   // (its barriers order the -1 fill above)
#pragma unroll
        for (int j = 0; j < kMaxKeysPerThread; ++j) {
            const int k = tid * kpt + j;
            end += sk[j];
            if (j < kpt && k < K) {
                const int base = end - tk[j];
                gb[k] = base + hv[j];
                if (blockIdx.x == 0) {
                    a.chainOff[k] = base;
                    a.chainEnd[k] = end;
                    // the chain is summed by a wave when it holds a multiple of L; the first such
                    // position names it in the wave list
                    const int L = 1 << a.longShift;
                    const int m0 = (base + L - 1) >> a.longShift;
                    if (tk[j] > 0 && m0 * L < end) a.longList[m0] = k;
                }
            }
        }
        if (blockIdx.x == 0 && tid == 0) *a.mCount = all;
    }
    __syncthreads();

    // rank the item's samples by key in sample order
    const int bits = K > 1 ? 32 - __clz(K - 1) : 1;
    const uint64_t below = lanes_below();
    int key[C], pos[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int l = lraw[c];
        int k = l >= 0 ? (int)(l - sd.sOff) : -1;
        if (k >= K) k = -1;
        key[c] = k;
        const bool valid = k >= 0;
        uint64_t m = __ballot(valid);
        for (int bit = 0; bit < bits; ++bit) {
            const bool bt = (k >> bit) & 1;
            const uint64_t bb = __ballot(valid && bt);
            m &= bt ? bb : ~bb;
        }
        int ps = 0;
        if (valid) ps = cnt[wave * keyCap + k] + __popcll(m & below);
        // the lowest lane of each group of equal keys advances that key's counter; within a
        // wave LDS operations complete in issue order, so the reads above see it before
        if (valid && (m & below) == 0) cnt[wave * keyCap + k] += __popcll(m);
        pos[c] = ps;
    }
    __syncthreads();

    // per key: the waves' bases inside the block, the block's count, block-local key bases
    int blockCount;
    {
        int bt[kMaxKeysPerThread];
        int sum = 0;
#pragma unroll
        for (int j = 0; j < kMaxKeysPerThread; ++j) {
            const int k = tid * kpt + j;
            int run = 0;
            if (j < kpt && k < K) {
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    const int t = cnt[w * keyCap + k];
                    cnt[w * keyCap + k] = run;
                    run += t;
                }
            }
            bt[j] = run;
            sum += run;
        }
        int ex = block_excl_scan<W>(sum, wsum, blockCount);
#pragma unroll
        for (int j = 0; j < kMaxKeysPerThread; ++j) {
            const int k = tid * kpt + j;
            if (j < kpt && k < K) lbase[k] = ex;
            ex += bt[j];
        }
    }
    __syncthreads();
    int dst[C], rk[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int k = key[c];
        dst[c] = 0;
        rk[c] = 0;
        if (k >= 0) {
            const int wb = cnt[wave * keyCap + k];
            dst[c] = gb[k] + wb + pos[c];
            rk[c] = lbase[k] + wb + pos[c];
        }
    }
    __syncthreads();   // the counters are dead: stage in block-local sorted order
    int* sDst = lds;
    T* sVal = reinterpret_cast<T*>(lds + E);
#pragma unroll
    for (int c = 0; c < C; ++c)
        if (key[c] >= 0) sDst[rk[c]] = dst[c];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (key[c] >= 0) sVal[rk[c]] = p[c][d];
        __syncthreads();
        T* __restrict__ out = a.sortedP + (int64_t)d * a.mStride;
        for (int r = tid; r < blockCount; r += TH) out[sDst[r]] = sVal[r];
        __syncthreads();
    }
}

inline int round_up_int(int v, int m) { return (v + m - 1) / m * m; }

template <typename T, int DIM, int W, int C>
void launch_scatter(const ChainArgs<T>& a, int numItems, int keyCap, hipStream_t st) {
    constexpr int E = W * kWave * C;
    constexpr int kStage = E + E * (int)(sizeof(T) / sizeof(int));
    const int region = std::max(W * keyCap, kStage);
    const size_t lds = sizeof(int) * ((size_t)region + 2 * (size_t)keyCap + W);
    hipLaunchKernelGGL((member_scatter_kernel<T, DIM, W, C>), dim3((unsigned)numItems), dim3(W * kWave), lds, st, a,
                       keyCap);
}

template <typename T, int DIM>
bool launch_chain_dim(const ChainArgs<T>& a, int numItems, int itemSize, int kb, hipStream_t st) {
    const int keyCap = round_up_int(std::max(kb, 1), 256);
    switch (itemSize) {
        case 128: launch_scatter<T, DIM, 2, 1>(a, numItems, keyCap, st); break;
        case 256: launch_scatter<T, DIM, 4, 1>(a, numItems, keyCap, st); break;
        case 512: launch_scatter<T, DIM, 4, 2>(a, numItems, keyCap, st); break;
        case 1024: launch_scatter<T, DIM, 4, 4>(a, numItems, keyCap, st); break;
        case 2048: launch_scatter<T, DIM, 4, 8>(a, numItems, keyCap, st); break;
        default: return false;
    }
    return true;
}

}  // namespace

bool chain_iteration_supported(int dim, int itemSize, int64_t kb) {
    const bool size = itemSize == 128 || itemSize == 256 || itemSize == 512 || itemSize == 1024 || itemSize == 2048;
    return dim >= 1 && dim <= kChainMaxDim && size && kb >= 1 && kb <= kChainMaxKeys;
}

template <typename T>
hipError_t launch_chain_order(const ChainArgs<T>& a, int dim, int numItems, int itemSize, int kb, hipStream_t st) {
    if (!chain_iteration_supported(dim, itemSize, kb) || a.keyBound != kb) return hipErrorInvalidValue;
    if (numItems <= 0) return hipSuccess;
    hipLaunchKernelGGL(key_prefix_kernel, dim3((unsigned)((kb + 3) / 4)), dim3(256), 0, st, a.hist, a.histStride,
                       numItems, a.segK, kb, a.totals, a.allDone);
    bool ok = false;
    switch (dim) {
        case 1: ok = launch_chain_dim<T, 1>(a, numItems, itemSize, kb, st); break;
        case 2: ok = launch_chain_dim<T, 2>(a, numItems, itemSize, kb, st); break;
        case 3: ok = launch_chain_dim<T, 3>(a, numItems, itemSize, kb, st); break;
        case 4: ok = launch_chain_dim<T, 4>(a, numItems, itemSize, kb, st); break;
        default: break;
    }
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template hipError_t launch_chain_order<float>(const ChainArgs<float>&, int, int, int, int, hipStream_t);
template hipError_t launch_chain_order<double>(const ChainArgs<double>&, int, int, int, int, hipStream_t);

}  // namespace rdhip
