// rd_chain.hip -- member ordering for evolve iterations of segments (whole clouds, the clouds
// of a batch, tiles) with at most kChainMaxKeys chosen points each, after the scan
// (rd_scan_impl.hpp) has written every sample's label and every block's label histogram.
//
// Reference semantics: rd/cpu/brute_force/evolve.hpp:32-77 (the members of every chosen point,
// summed in sample order).
//
//   key_prefix     -- one wave per (segment, key): the key's row of block counts becomes its
//                     exclusive prefix over the segment's blocks, and the key's member total.
//   member_scatter -- one block per scan item (same samples, same XCD-aware order): ranks the
//                     item's samples by key in sample order (ballot match per key bit, as in
//                     rd_sort.hip), reads their coordinates coalesced, and writes them in chain
//                     order through an LDS stage, so consecutive threads write consecutive
//                     addresses of each chain run.  A segment's chain k lies in its region at
//                     [chainOff, chainEnd), chainEnd the running sum of the totals rounded up to
//                     16 bytes.  Each segment's first block also writes the chain ranges and
//                     lists the long chains for the ordered sums (chain_sum_shift_kernel,
//                     rd_evolve.hip), which follow.
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
// per-key prefix over a segment's blocks: one wave per (segment, key) row
// ---------------------------------------------------------------------------------
// segment of a row: the last s with segRowOff[s] <= row
__device__ __forceinline__ int row_segment(const int* __restrict__ rowOff, int numSegs, int row) {
    int lo = 0, hi = numSegs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (rowOff[mid] <= row) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <typename T>
__global__ __launch_bounds__(256) void key_prefix_kernel(ChainArgs<T> a) {
    if (*a.allDone) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.longCount = 0;   // the scatter's long-chain list
    const int lane = lane_id();
    const int row = blockIdx.x * 4 + threadIdx.x / kWave;
    if (row >= a.totalRows) return;
    const int s = row_segment(a.segRowOff, a.numSegs, row);
    const int k = row - a.segRowOff[s];
    const SegDesc sd = a.segs[s];
    const int nb = (sd.nP + a.itemSize - 1) / a.itemSize;
    int* __restrict__ r = a.hist + a.segHistOff[s] + (int64_t)k * nb;
    constexpr int kPer = 16;   // one pass for up to 1024 blocks
    int carry = 0;
    for (int base = 0; base < nb; base += kWave * kPer) {
        const int i0 = base + lane * kPer;
        int v[kPer];
        int sum = 0;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            v[j] = i0 + j < nb ? r[i0 + j] : 0;
            sum += v[j];
        }
        // rows of converged segments and of keys past the live count hold stale counts: loaded
        // (in bounds, in parallel with the flags' loads) but neither rewritten nor totalled
        if (a.segDone[s] || k >= min(a.segK[s], a.segKeyCap[s])) {
            if (lane == 0) a.totals[row] = 0;
            return;
        }
        int incl = sum;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        int run = carry + incl - sum;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            if (i0 + j < nb) r[i0 + j] = run;
            run += v[j];
        }
        carry += __shfl(incl, kWave - 1);
    }
    if (lane == 0) a.totals[row] = carry;
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
    const int nItems = a.numItemsDev ? *a.numItemsDev : (int)gridDim.x;
    if ((int)blockIdx.x >= nItems) return;
    const WorkItem it = a.items[xcd_item(blockIdx.x, nItems)];
    const int s = it.seg;
    if (a.segDone[s]) return;
    const SegDesc sd = a.segs[s];
    const int K = min(min(a.segK[s], a.segKeyCap[s]), keyCap);
    const int col = it.start / a.itemSize;
    const int nb = (sd.nP + a.itemSize - 1) / a.itemSize;
    const int rowOff = a.segRowOff[s];
    const int64_t memberOff = a.segMemberOff[s];
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
    const int* __restrict__ hcol = a.hist + a.segHistOff[s] + col;
    int tk[kMaxKeysPerThread], hv[kMaxKeysPerThread];
#pragma unroll
    for (int j = 0; j < kMaxKeysPerThread; ++j) {
        const int k = tid * kpt + j;
        const bool live = j < kpt && k < K;
        tk[j] = live ? a.totals[rowOff + k] : 0;
        hv[j] = live ? hcol[(int64_t)k * nb] : 0;
    }

    for (int v = tid; v < W * keyCap; v += TH) cnt[v] = 0;
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
        int end = block_excl_scan<W>(sum, wsum, all);
#pragma unroll
        for (int j = 0; j < kMaxKeysPerThread; ++j) {
            const int k = tid * kpt + j;
            end += sk[j];
            if (j < kpt && k < K) {
                const int64_t base = memberOff + end - tk[j];
                gb[k] = (int)(base + hv[j]);
                if (col == 0) {
                    const int64_t slot = sd.sOff + k;
                    a.chainOff[slot] = (int)base;
                    a.chainEnd[slot] = (int)(memberOff + end);
                    // long chains take a place in the wave list (which place does not matter:
                    // a chain's sum is the same on either path)
                    if (tk[j] >= a.longMin) {
                        const int idx = atomicAdd(a.longCount, 1);
                        a.longIdx[slot] = idx;
                        if (idx < a.longCap) a.longList[idx] = (int)slot;
                    }
                }
            }
        }
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
bool launch_chain_dim(const ChainArgs<T>& a, int numItems, int keyCapMax, hipStream_t st) {
    const int keyCap = round_up_int(std::max(keyCapMax, 1), 256);
    switch (a.itemSize) {
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

bool chain_iteration_supported(int dim, int itemSize, int64_t keyCapMax) {
    const bool size = itemSize == 128 || itemSize == 256 || itemSize == 512 || itemSize == 1024 || itemSize == 2048;
    return dim >= 1 && dim <= kChainMaxDim && size && keyCapMax >= 1 && keyCapMax <= kChainMaxKeys;
}

template <typename T>
hipError_t launch_chain_order(const ChainArgs<T>& a, int dim, int numItems, int keyCapMax, hipStream_t st) {
    if (!chain_iteration_supported(dim, a.itemSize, keyCapMax)) return hipErrorInvalidValue;
    if (numItems <= 0 || a.totalRows <= 0) return hipSuccess;
    hipLaunchKernelGGL((key_prefix_kernel<T>), dim3((unsigned)((a.totalRows + 3) / 4)), dim3(256), 0, st, a);
    bool ok = false;
    RD_DISPATCH_DIM(dim, { ok = launch_chain_dim<T, DIM>(a, numItems, keyCapMax, st); });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

template hipError_t launch_chain_order<float>(const ChainArgs<float>&, int, int, int, hipStream_t);
template hipError_t launch_chain_order<double>(const ChainArgs<double>&, int, int, int, hipStream_t);

}  // namespace rdhip
