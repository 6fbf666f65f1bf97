// rd_local.hip -- evolve of small segments (tiles, small clouds of a batch), one
// workgroup per segment, every iteration inside the workgroup.
//
// The batch-wide iteration (rd_evolve.hip) pays about fifteen launches and a device-wide
// sort per iteration, however little work is left; a segment of up to a few ten thousand
// samples iterates much faster on one CU with its chosen points, running sums and member
// lists in LDS.  Per iteration the samples are taken in chunks of kLocalChunk, in order:
//   1. labels: nearest chosen point of every sample of the chunk (rd_nearest.hpp, chosen
//      points read from LDS), stored as int16 in LDS;
//   2. stable counting sort of the chunk by label: wave w owns a contiguous range of the
//      chunk; it counts the labels of its range (ballot match, one LDS add per distinct
//      label and 64 samples), a block scan turns the counts into per-(wave, label) bases,
//      and the wave walks its range again placing every sample at base + rank among the
//      equal labels of its 64-sample chunk -- so each chosen point's members are listed in
//      sample order;
//   3. ordered sums: one 16-lane row per (coordinate, chosen point) continues that
//      point's running sum with the chunk's members in sample order (rd_rowsum.hpp: the
//      reference's serial sum);
// then 4. the shift rule updates the chosen points in LDS, and the workgroup repeats until
// no coordinate moved (or the iteration cap).  Results are bit-identical to the batch-wide
// path: same labels, same member order, same additions.  Reference:
// rd/cpu/brute_force/evolve.hpp:32-77 (labels, sums), 129-179 (shift), 237-459 (loop).
#include "rd_local.hpp"
#include "rd_nearest.hpp"
#include "rd_rowsum.hpp"

#include <algorithm>

namespace rdhip {

namespace {

constexpr int kSmallInts = 16;   // wave totals of the block scan + two moved flags
constexpr int kLocalRing = 4;    // steps in flight in the ordered sums (members are L2-resident)

__host__ __device__ inline size_t align16(size_t v) { return (v + 15) / 16 * 16; }

// LDS layout: chosen coordinates [dim][kStride] | running sums [dim][kStride] | total
// counts, chunk counts, chunk offsets [kStride] each | per-wave bases [waves][kStride] |
// small ints | labels int16 [chunk] | order uint16 [chunk]
__host__ __device__ inline size_t lds_bytes(int dim, size_t vb, int kStride, int chunk) {
    return 2 * align16((size_t)dim * kStride * vb) + 4 * (size_t)kStride * (3 + kLocalWaves) + 4 * kSmallInts +
           2 * align16(2 * (size_t)chunk);
}

// LDS arrays of the segment-local kernels
template <typename T>
struct LocalLds {
    T* sS;
    T* run;
    int* cntT;
    int* cnt;
    int* off;
    int* wbase;   // [wave][label]
    int* small;   // [0, waves): wave totals; [14], [15]: moved flags
    short* lab;
    unsigned short* order;
    int ks;
};

template <typename T, int DIM>
__device__ __forceinline__ LocalLds<T> local_lds(char* smem, int ks, int chunk) {
    LocalLds<T> L;
    L.ks = ks;
    L.sS = reinterpret_cast<T*>(smem);
    L.run = reinterpret_cast<T*>(smem + align16((size_t)DIM * ks * sizeof(T)));
    L.cntT = reinterpret_cast<int*>(smem + 2 * align16((size_t)DIM * ks * sizeof(T)));
    L.cnt = L.cntT + ks;
    L.off = L.cnt + ks;
    L.wbase = L.off + ks;
    L.small = L.wbase + kLocalWaves * ks;
    L.lab = reinterpret_cast<short*>(L.small + kSmallInts);
    L.order = reinterpret_cast<unsigned short*>(reinterpret_cast<char*>(L.lab) + align16(2 * (size_t)chunk));
    return L;
}

// With the labels of samples c0 .. c0+cn-1 of the segment in lab[0 .. cn) (local chosen
// index or -1) and wbase zeroed (barrier passed): stable counting sort of the chunk by
// label into order, chunk counts/offsets into cnt/off, totals added to cntT, then every
// (coordinate, chosen point) running sum in run continued with the chunk's members in
// sample order.  Ends with a barrier.
template <typename T, int DIM>
__device__ __forceinline__ void sort_and_fold_chunk(const LocalArgs<T>& a, const LocalLds<T>& Ld, const SegDesc& sd,
                                                    int K, int c0, int cn, int bits) {
    const int ks = Ld.ks;
    short* lab = Ld.lab;
    int* wbase = Ld.wbase;
    int* small = Ld.small;
    int* cnt = Ld.cnt;
    int* off = Ld.off;
    int* cntT = Ld.cntT;
    T* run = Ld.run;
    unsigned short* order = Ld.order;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const uint64_t below = lanes_below();
    const int64_t ps = a.pStride;
    // ballot match: lanes whose label equals this lane's (valid lanes only)
    auto match = [&](int Lb) {
        const bool v = Lb >= 0;
        uint64_t m = __ballot(v);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (Lb >> b) & 1;
            const uint64_t bb = __ballot(v && bit);
            m &= bit ? bb : ~bb;
        }
        return m;
    };
    // 2a. per-wave label counts over the wave's range of the chunk
    const int range = ((cn + kLocalWaves - 1) / kLocalWaves + kWave - 1) / kWave * kWave;
    const int r0 = min(cn, wave * range), r1 = min(cn, r0 + range);
    for (int q0 = r0; q0 < r1; q0 += kWave) {
        const int i = q0 + lane;
        const int L = i < r1 ? lab[i] : -1;
        const uint64_t m = match(L);
        if (L >= 0 && (m & below) == 0) wbase[wave * ks + L] += __popcll(m);
    }
    __syncthreads();

    // 2b. chunk counts, offsets (block exclusive scan, two labels per thread), wave bases
    {
        const int k0 = 2 * tid, k1 = 2 * tid + 1;
        int t0 = 0, t1 = 0;
        if (k0 < K)
            for (int w = 0; w < kLocalWaves; ++w) t0 += wbase[w * ks + k0];
        if (k1 < K)
            for (int w = 0; w < kLocalWaves; ++w) t1 += wbase[w * ks + k1];
        const int s = t0 + t1;
        int x = s;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == kWave - 1) small[wave] = x;
        __syncthreads();
        int pre = 0;
        for (int w = 0; w < wave; ++w) pre += small[w];
        const int ex = pre + x - s;
        if (k0 < K) {
            off[k0] = ex;
            cnt[k0] = t0;
            cntT[k0] += t0;
            int r = ex;
            for (int w = 0; w < kLocalWaves; ++w) {
                const int c = wbase[w * ks + k0];
                wbase[w * ks + k0] = r;
                r += c;
            }
        }
        if (k1 < K) {
            off[k1] = ex + t0;
            cnt[k1] = t1;
            cntT[k1] += t1;
            int r = ex + t0;
            for (int w = 0; w < kLocalWaves; ++w) {
                const int c = wbase[w * ks + k1];
                wbase[w * ks + k1] = r;
                r += c;
            }
        }
    }
    __syncthreads();

    // 2c. placement in sample order.  Within a wave LDS operations complete in
    // issue order, so every lane reads its label's base before the leader advances it.
    for (int q0 = r0; q0 < r1; q0 += kWave) {
        const int i = q0 + lane;
        const int L = i < r1 ? lab[i] : -1;
        const uint64_t m = match(L);
        if (L >= 0) order[wbase[wave * ks + L] + __popcll(m & below)] = (unsigned short)i;
        if (L >= 0 && (m & below) == 0) wbase[wave * ks + L] += __popcll(m);
    }
    __syncthreads();

    // 3. running sums, rows of 16 lanes over (coordinate, chosen point)
    for (int rb = 0; rb < DIM * K; rb += kLocalThreads / 16) {
        const int row = rb + tid / 16;
        const bool act = row < DIM * K;
        const int d = act ? row / K : 0;
        const int k = act ? row - d * K : 0;
        const int c = act ? cnt[k] : 0;
        const int o = act ? off[k] : 0;
        const T* __restrict__ Pd = a.P + (int64_t)d * ps;
        const T sum = row_ordered_sum<T, kLocalRing>(
            wave_row_max(c),
            [&](int i) {
                const int pos = c > 0 ? (int)order[o + min(i, c - 1)] : 0;
                const T x = Pd[point_of(a.pIndex, sd.pOff + c0 + pos)];
                return i < c ? x : T(0);
            },
            act ? run[d * ks + k] : T(0));
        if (act && (tid & 15) == 0) run[d * ks + k] = sum;
    }
    __syncthreads();
}

template <typename T, int DIM, int PPT>
__global__ __launch_bounds__(kLocalThreads) void local_evolve_kernel(LocalArgs<T> a) {
    extern __shared__ __align__(16) char smem[];
    const int ks = a.kStride;
    const LocalLds<T> Ld = local_lds<T, DIM>(smem, ks, a.pCap);
    const int chunk = a.pCap;
    T* sS = Ld.sS;
    T* run = Ld.run;
    int* cntT = Ld.cntT;
    int* wbase = Ld.wbase;
    short* lab = Ld.lab;
    int* moved = Ld.small + 14;

    const int seg = a.list[blockIdx.x];
    const SegDesc sd = a.segs[seg];
    const int K = a.segK[seg];
    const int nP = sd.nP;
    const T rSqr = a.segRSqr[seg];
    const int tid = threadIdx.x;
    const int64_t ps = a.pStride;

    // chosen points into LDS, the far-away sentinel after K up to the end of the last tile
    const int kt = (K + kChosenTile - 1) / kChosenTile * kChosenTile;
    for (int i = tid; i < DIM * kt; i += kLocalThreads) {
        const int d = i / kt, k = i - d * kt;
        sS[d * ks + k] = k < K ? a.S[d * a.sStride + sd.sOff + k] : FpTraits<T>::sentinel;
    }
    if (tid < 2) moved[tid] = 0;
    __syncthreads();

    int iter = 0;
    if (K > 0 && nP > 0) {
        const int bits = K > 1 ? 32 - __clz(K - 1) : 1;
        for (;;) {
            for (int i = tid; i < DIM * K; i += kLocalThreads) run[(i / K) * ks + i % K] = T(0);
            for (int i = tid; i < K; i += kLocalThreads) cntT[i] = 0;
            for (int c0 = 0; c0 < nP; c0 += chunk) {
                const int cn = min(chunk, nP - c0);
                // 1. labels of the chunk
                for (int base = 0; base < cn; base += kLocalThreads * PPT) {
                    T p[PPT][DIM];
                    bool valid[PPT];
#pragma unroll
                    for (int j = 0; j < PPT; ++j) {
                        const int pos = base + tid + j * kLocalThreads;
                        valid[j] = pos < cn;
                        const int64_t src = valid[j] ? point_of(a.pIndex, sd.pOff + c0 + pos) : 0;
#pragma unroll
                        for (int d = 0; d < DIM; ++d) p[j][d] = valid[j] ? a.P[d * ps + src] : T(0);
                    }
                    int label[PPT];
                    nearest_labels<T, DIM, PPT>(p, valid, K, sS, ks, rSqr, label);
#pragma unroll
                    for (int j = 0; j < PPT; ++j)
                        if (valid[j]) lab[base + tid + j * kLocalThreads] = (short)label[j];
                }
                for (int i = tid; i < kLocalWaves * K; i += kLocalThreads) wbase[(i / K) * ks + i % K] = 0;
                __syncthreads();
                if (tid == 0 && c0 == 0) moved[(iter + 1) & 1] = 0;

                sort_and_fold_chunk<T, DIM>(a, Ld, sd, K, c0, cn, bits);
            }

            // 4. shift toward the mass centres
            for (int i = tid; i < DIM * K; i += kLocalThreads) {
                const int d = i / K, k = i - d * K;
                const int c = cntT[k];
                if (c == 0) continue;
                const T fc = (T)c;
                const T mc = run[d * ks + k] / fc;
                T thr = T(2) * fabs(mc);
                thr = thr * fc;
                thr = thr * FpTraits<T>::eps;
                T& s = sS[d * ks + k];
                if (fabs(mc - s) > thr) {
                    s = mc;
                    moved[iter & 1] = 1;
                }
            }
            __syncthreads();
            const bool any = moved[iter & 1] != 0;
            ++iter;
            if (!any || (a.maxIter > 0 && iter >= a.maxIter)) break;
            if (iter >= a.guard) {
                if (tid == 0) *a.error = 1;
                break;
            }
        }
    } else {
        iter = 1;   // the reference loop runs one pass that moves nothing
    }

    for (int i = tid; i < DIM * K; i += kLocalThreads) {
        const int d = i / K, k = i - d * K;
        a.S[d * a.sStride + sd.sOff + k] = sS[d * ks + k];
    }
    if (tid == 0) {
        a.segIter[seg] = iter;
        atomicMax(a.maxIters, iter);
    }
}

// One evolve iteration of every segment with its labels taken from the batch-wide scan
// (global chosen slots, or -1): the counting sort and the ordered sums of the local kernel
// in one workgroup per segment, then the shift straight on the chosen points in HBM and the
// segment's moved flag.  Replaces the batch-wide sort, member gather and sums for batches
// of small segments.
template <typename T, int DIM>
__global__ __launch_bounds__(kLocalThreads) void segment_sums_kernel(LocalArgs<T> a) {
    if (*a.allDone) return;
    const int seg = blockIdx.x;
    if (a.segDone[seg]) return;
    extern __shared__ __align__(16) char smem[];
    const int ks = a.kStride;
    const LocalLds<T> Ld = local_lds<T, DIM>(smem, ks, a.pCap);
    const SegDesc sd = a.segs[seg];
    const int K = a.segK[seg];
    const int nP = sd.nP;
    if (K <= 0 || nP <= 0) return;   // no member: nothing moves
    const int tid = threadIdx.x;
    const int chunk = a.pCap;
    const int bits = K > 1 ? 32 - __clz(K - 1) : 1;
    for (int i = tid; i < DIM * K; i += kLocalThreads) Ld.run[(i / K) * ks + i % K] = T(0);
    for (int i = tid; i < K; i += kLocalThreads) Ld.cntT[i] = 0;
    for (int c0 = 0; c0 < nP; c0 += chunk) {
        const int cn = min(chunk, nP - c0);
        for (int i = tid; i < cn; i += kLocalThreads) {
            const int g = a.labelsIn[sd.pOff + c0 + i];
            Ld.lab[i] = (short)(g < 0 ? -1 : (int)(g - sd.sOff));
        }
        for (int i = tid; i < kLocalWaves * K; i += kLocalThreads) Ld.wbase[(i / K) * ks + i % K] = 0;
        __syncthreads();
        sort_and_fold_chunk<T, DIM>(a, Ld, sd, K, c0, cn, bits);
    }
    bool movedAny = false;
    for (int i = tid; i < DIM * K; i += kLocalThreads) {
        const int d = i / K, k = i - d * K;
        const int c = Ld.cntT[k];
        if (c == 0) continue;
        const T fc = (T)c;
        const T mc = Ld.run[d * ks + k] / fc;
        T thr = T(2) * fabs(mc);
        thr = thr * fc;
        thr = thr * FpTraits<T>::eps;
        T& s = a.S[d * a.sStride + sd.sOff + k];
        if (fabs(mc - s) > thr) {
            s = mc;
            movedAny = true;
        }
    }
    if (movedAny) a.segMoved[seg] = 1;
}

template <typename T, int DIM>
hipError_t launch_segment_dim(const LocalArgs<T>& a, int numSegs, hipStream_t st) {
    const size_t bytes = lds_bytes(DIM, sizeof(T), a.kStride, a.pCap);
    static bool attrSet = false;
    if (!attrSet) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&segment_sums_kernel<T, DIM>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLocalLdsBudget);
        if (e != hipSuccess) return e;
        attrSet = true;
    }
    hipLaunchKernelGGL((segment_sums_kernel<T, DIM>), dim3(numSegs), dim3(kLocalThreads), bytes, st, a);
    return hipGetLastError();
}

template <typename T, int DIM>
hipError_t launch_dim(const LocalArgs<T>& a, hipStream_t st) {
    constexpr int PPT = DIM <= 4 ? 2 : 1;
    const size_t bytes = lds_bytes(DIM, sizeof(T), a.kStride, a.pCap);
    static bool attrSet = false;
    if (!attrSet) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&local_evolve_kernel<T, DIM, PPT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLocalLdsBudget);
        if (e != hipSuccess) return e;
        attrSet = true;
    }
    hipLaunchKernelGGL((local_evolve_kernel<T, DIM, PPT>), dim3(a.count), dim3(kLocalThreads), bytes, st, a);
    return hipGetLastError();
}

}  // namespace

size_t local_lds_bytes(int dim, size_t valueBytes, int kStride, int pCap) {
    return lds_bytes(dim, valueBytes, kStride, pCap);
}

template <typename T>
hipError_t launch_local_evolve(const LocalArgs<T>& a, int dim, hipStream_t st) {
    if (a.count <= 0) return hipSuccess;
    if (local_lds_bytes(dim, sizeof(T), a.kStride, a.pCap) > kLocalLdsBudget || a.pCap > kLocalChunk ||
        a.kStride > kLocalMaxK + kChosenTile)
        return hipErrorInvalidValue;
    hipError_t e = hipErrorInvalidValue;
    RD_DISPATCH_DIM(dim, { e = launch_dim<T, DIM>(a, st); });
    return e;
}

template <typename T>
hipError_t launch_segment_sums(const LocalArgs<T>& a, int dim, int numSegs, hipStream_t st) {
    if (numSegs <= 0) return hipSuccess;
    if (local_lds_bytes(dim, sizeof(T), a.kStride, a.pCap) > kLocalLdsBudget || a.pCap > kLocalChunk ||
        a.kStride > kLocalMaxK + kChosenTile)
        return hipErrorInvalidValue;
    hipError_t e = hipErrorInvalidValue;
    RD_DISPATCH_DIM(dim, { e = launch_segment_dim<T, DIM>(a, numSegs, st); });
    return e;
}

template hipError_t launch_segment_sums<float>(const LocalArgs<float>&, int, int, hipStream_t);
template hipError_t launch_segment_sums<double>(const LocalArgs<double>&, int, int, hipStream_t);
template hipError_t launch_local_evolve<float>(const LocalArgs<float>&, int, hipStream_t);
template hipError_t launch_local_evolve<double>(const LocalArgs<double>&, int, hipStream_t);

}  // namespace rdhip
