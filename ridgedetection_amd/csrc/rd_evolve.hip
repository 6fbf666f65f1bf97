// rd_evolve.hip -- one evolve iteration for a batch of segments, on gfx950.
//
// Reference semantics (one iteration, rd/cpu/brute_force/evolve.hpp:32-77, 129-179):
//   for every sample n (in order): k* = first argmin_k d2(S_k, P_n); if d2 <= r*r then
//   count[k*]++ and sum[k*] += P_n; then for every chosen k and coordinate d with
//   count > 0: mc = sum / count; if |mc - S| > ((2|mc|) count) eps: S = mc, flag = 1.
//
// Device pipeline per iteration:
//   1. evolve_scan (rd_scan_impl.hpp) -- nearest chosen point of every sample
//      (rd_nearest.hpp): the K chosen
//      points of a segment stream through SGPRs (uniform scalar loads) in tiles of 16;
//      every lane keeps 4 samples; screened distances with an exact decision.  No MFMA:
//      the d <= 12 inner product feeds an argmin.
//   2. stable sort of the samples by label (rd_sort.hip); its last pass writes the
//      members' coordinates in chain order (every chosen point's members contiguous, in
//      sample order).
//   3. chain_bounds -- start/end of every chosen point's member run.
//   4. chain_sum_shift -- one 16-lane row per (coordinate, chosen point) folds the members
//      into the running sum in sample order (the reference's serial summation, so sums are
//      bit-identical), applies the shift rule and raises the segment's "moved" flag.
//      Chains holding a sorted position that is a multiple of L (every chain of at least L
//      members; L adapts so that at most kLongWaves entries are needed), or listed by length,
//      go to a block of four waves each instead, in the same launch: each wave sums a quarter
//      of the chain from 64 speculative starts at once, one per lane, and the quarter whose
//      start matches the exact sum before it is kept (chain_sum_parts).
//   5. evolve_finalize -- per-segment iteration count and convergence; global all-done.
#include "rd_evolve.hpp"
#include "rd_rowsum.hpp"

namespace rdhip {

// ---------------------------------------------------------------------------------
// 3. member runs of every chosen slot in the sorted label array
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void chain_bounds_kernel(const int* __restrict__ keys,
                                                           const int64_t* __restrict__ mPtr,
                                                           int* __restrict__ chainOff,
                                                           int* __restrict__ chainEnd,
                                                           int* __restrict__ longList, int longShift,
                                                           const int* __restrict__ allDone) {
    if (*allDone) return;
    const int64_t M = *mPtr;
    const int64_t L = int64_t(1) << longShift;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < M; i += (int64_t)gridDim.x * 256) {
        int k = keys[i];
        if (i == 0 || keys[i - 1] != k) chainOff[k] = (int)i;
        if (i == M - 1 || keys[i + 1] != k) chainEnd[k] = (int)(i + 1);
        // i = m L is the first multiple of L inside its chain iff i - L lies outside it
        if (longList && (i & (L - 1)) == 0) longList[i >> longShift] = (i < L || keys[i - L] != k) ? k : -1;
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
// ---------------------------------------------------------------------------------
// 4b'. ordered sums through LDS: a group of kLanes lanes (a whole wave, or a 16-lane row)
// sums one chain.  Tiles of kLanes * kLongPer members are loaded coalesced one tile before
// their turn (the loads of tile t + 1 are in flight while tile t is summed), staged in the
// group's LDS slice and added in member order by every lane of the group from broadcast
// LDS reads, so each member costs one plain dependent add (6.25 cycles, against 13.25 for
// an add fed by a DPP row broadcast).  Members past the end of the chain are staged as +0
// (see above).  The four rows of a wave read four slices whose starts differ by 16 bytes
// modulo the bank width, so their broadcasts do not collide.
// ---------------------------------------------------------------------------------
constexpr int kLongPer = 16;
constexpr int kLongTile = kLongPer * kWave;
constexpr int kRowTile = kLongPer * 16;

template <typename T>
struct WideOf;
template <>
struct WideOf<float> { using type = float4; static constexpr int n = 4; };
template <>
struct WideOf<double> { using type = double2; static constexpr int n = 2; };

// LDS per wave: one wave slice, or four row slices padded by one wide read each
template <typename T>
constexpr int wave_lds_elems() {
    return kLongTile > 4 * (kRowTile + WideOf<T>::n) ? kLongTile : 4 * (kRowTile + WideOf<T>::n);
}

// sum + v.x + v.y (+ v.z + v.w), in that order
__device__ __forceinline__ float add_wide(float sum, float4 v) {
    sum = sum + v.x;
    sum = sum + v.y;
    sum = sum + v.z;
    return sum + v.w;
}
__device__ __forceinline__ double add_wide(double sum, double2 v) {
    sum = sum + v.x;
    return sum + v.y;
}

// orders a wave's LDS accesses (the slices are private to the wave)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Serial sum start + src[0] + src[1] + ... + src[cnt - 1] by a group of kLanes lanes (every
// lane of the wave takes part; `steps` is the wave-uniform largest cnt of its groups); tile is
// the group's LDS slice.  `start` may differ between lanes (each lane then returns its own sum)
// and must not be -0 (the zeros of the first pass and of the padding leave any other value as
// it is).  load(0) may be issued early (the first tile's members are then in reg) and the sum
// run later.
template <typename T, int kLanes>
struct OrderedRun {
    static constexpr int kTile = kLanes * kLongPer;
    const T* __restrict__ src;
    int cnt, last, lane;
    T reg[kLongPer];

    __device__ __forceinline__ OrderedRun(const T* __restrict__ s, int n)
        : src(s), cnt(n), last(max(n - 1, 0)), lane((int)(threadIdx.x & (kLanes - 1))) {}

    // tile t: index clamped into the chain and masked (tile t + 1 is loaded after tile t is
    // staged; its loads are not waited for before tile t's sum)
    __device__ __forceinline__ void load(int t) {
#pragma unroll
        for (int c = 0; c < kLongPer; ++c) {
            const int i = t * kTile + c * kLanes + lane;
            const T x = src[min(i, last)];
            reg[c] = i < cnt ? x : T(0);
        }
    }

    // the sum, with tile 0 already in reg
    __device__ __forceinline__ T run(int steps, T* tile, T start) {
        using W = typename WideOf<T>::type;
        constexpr int kW = WideOf<T>::n;
        const int nTiles = (steps + kTile - 1) / kTile;
        T sum = start;
        for (int t = 0; t < nTiles; ++t) {
            wave_lds_sync();
#pragma unroll
            for (int c = 0; c < kLongPer; ++c) tile[c * kLanes + lane] = reg[c];
            wave_lds_sync();
            load(t + 1);
            // groups of kG wide reads: group g + 1 is read while group g is added (the first
            // pass adds a group of zeros, which leaves the sum unchanged); the scheduling
            // barriers keep the reads a whole group ahead of their adds
            const W* w = reinterpret_cast<const W*>(tile);
            constexpr int kG = 8, kNQ = kTile / kW;
            static_assert(kNQ % kG == 0, "tile is not a whole number of read groups");
            W cur[kG];
#pragma unroll
            for (int g = 0; g < kG; ++g) cur[g] = W{};
#pragma unroll 2
            for (int q0 = 0; q0 < kNQ; q0 += kG) {
                W nxt[kG];
#pragma unroll
                for (int g = 0; g < kG; ++g) nxt[g] = w[q0 + g];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int g = 0; g < kG; ++g) sum = add_wide(sum, cur[g]);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int g = 0; g < kG; ++g) cur[g] = nxt[g];
            }
#pragma unroll
            for (int g = 0; g < kG; ++g) sum = add_wide(sum, cur[g]);
        }
        return sum;
    }
};

template <typename T, int kLanes>
__device__ __forceinline__ T group_ordered_sum(const T* __restrict__ src, int cnt, int steps, T* tile, T start) {
    OrderedRun<T, kLanes> r(src, cnt);
    r.load(0);
    return r.run(steps, tile, start);
}

// ---------------------------------------------------------------------------------
// 4b''. tolerance mode: parallel sums in a wider type.  Lane l of a group of kLanes lanes
// accumulates members l, l + kLanes, ... into four interleaved accumulators (so the adds do
// not wait on each other), combined in a fixed order and then across the group by a fixed
// xor butterfly: the result does not depend on scheduling.  Float data accumulates in
// double: the error of a sum of m members is at most about m 2^-53 times the sum of their
// magnitudes, below one float unit of that magnitude for m < 2^29.  Double data accumulates in
// double-double (Knuth's TwoSum), accurate to a few units of the last place of a double
// whatever the chain length.
// ---------------------------------------------------------------------------------
struct DD {
    double hi, lo;
};
__device__ __forceinline__ DD dd_add(DD a, double b) {
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double err = (a.hi - (s - bb)) + (b - bb);
    return DD{s, a.lo + err};
}
__device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD r = dd_add(a, b.hi);
    r.lo += b.lo;
    return r;
}
template <typename T>
struct WideSum;
template <>
struct WideSum<float> {
    using Acc = double;
    static __device__ __forceinline__ Acc zero() { return 0.0; }
    static __device__ __forceinline__ Acc add(Acc a, float x) { return a + (double)x; }
    static __device__ __forceinline__ Acc add(Acc a, Acc b) { return a + b; }
    static __device__ __forceinline__ Acc xorShfl(Acc a, int m) { return __shfl_xor(a, m); }
    static __device__ __forceinline__ double value(Acc a) { return a; }
};
template <>
struct WideSum<double> {
    using Acc = DD;
    static __device__ __forceinline__ Acc zero() { return DD{0.0, 0.0}; }
    static __device__ __forceinline__ Acc add(Acc a, double x) { return dd_add(a, x); }
    static __device__ __forceinline__ Acc add(Acc a, Acc b) { return dd_add(a, b); }
    static __device__ __forceinline__ Acc xorShfl(Acc a, int m) { return DD{__shfl_xor(a.hi, m), __shfl_xor(a.lo, m)}; }
    static __device__ __forceinline__ double value(Acc a) { return a.hi + a.lo; }
};

// Sum of src[0 .. cnt) by a group of kLanes lanes (all lanes of the group return it).
template <typename T, int kLanes>
__device__ __forceinline__ typename WideSum<T>::Acc group_parallel_sum(const T* __restrict__ src, int cnt) {
    using W = WideSum<T>;
    const int lane = threadIdx.x & (kLanes - 1);
    typename W::Acc acc[4] = {W::zero(), W::zero(), W::zero(), W::zero()};
    int i = lane;
    for (; i + 3 * kLanes < cnt; i += 4 * kLanes) {
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = W::add(acc[u], src[i + u * kLanes]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (i + u * kLanes < cnt) acc[u] = W::add(acc[u], src[i + u * kLanes]);
    typename W::Acc s = W::add(W::add(acc[0], acc[1]), W::add(acc[2], acc[3]));
#pragma unroll
    for (int m = kLanes / 2; m >= 1; m /= 2) s = W::add(s, W::xorShfl(s, m));
    return s;
}

// The shift rule of shift_coordinate with the mean taken from a wide sum.
template <typename T>
__device__ __forceinline__ void shift_coordinate_wide(const ShiftArgs<T>& a, int d, int64_t slot, int seg, int cnt,
                                                      double wide) {
    if (a.outCounts && d == 0) a.outCounts[slot] = cnt;
    if (a.outSums) a.outSums[d * a.sStride + slot] = (T)wide;
    bool moved = false;
    if (cnt != 0) {
        const T fc = (T)cnt;
        constexpr T eps = FpTraits<T>::eps;
        const T mc = (T)(wide / (double)cnt);
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

// 16-lane row path: one row per (coordinate, chosen slot) of chains not on the long list
template <typename T, int DIM>
__device__ __forceinline__ void chain_sum_rows(const ShiftArgs<T>& a, int64_t blk, T* waveLds) {
    constexpr int kRow = 16;
    const int t = threadIdx.x & (kRow - 1);
    const int64_t row = blk * (256 / kRow) + threadIdx.x / kRow;
    const int d = (int)(row / max<int64_t>(a.totalSlots, 1));
    const int64_t rs = row - (int64_t)d * a.totalSlots;
    const int64_t slot = (a.rowSlots && d < DIM) ? (int64_t)a.rowSlots[rs] : rs;
    const int64_t chain = a.denseChains ? rs : slot;
    int seg = -1, off = 0, cnt = 0;
    if (d < DIM) {
        // the slot's segment, range and list index are loaded together (in bounds for every slot;
        // used only once the slot is known to be live)
        seg = a.slotSeg[slot];
        const int o = a.chainOff[chain];
        const int ce = a.chainEnd[chain];
        const int li = a.longCount ? a.longIdx[chain] : 0;
        if (seg >= 0 && !a.segDone[seg] && (int)(slot - a.segs[seg].sOff) < a.segK[seg]) {
            off = o;
            cnt = ce - o;
            // a chain holding a multiple of L (or, by length, one of the listed long chains) is
            // summed by a block (chain_sum_parts)
            const int64_t s0 = off, e1 = (int64_t)off + cnt - 1;
            const bool waveSummed =
                a.longCount ? (cnt >= a.longMin && li < a.longCap)
                            : (a.longList && cnt > 0 &&
                               (e1 >> a.longShift) >= ((s0 + (int64_t(1) << a.longShift) - 1) >> a.longShift));
            if (waveSummed) {
                seg = -1;
                cnt = 0;
            }
        } else {
            seg = -1;
        }
    }
    // the wave steps through the longest of its four chains; every lane takes part in the
    // staging.  Loads are unconditional (index clamped into the chain, or element 0 of the
    // buffer for an empty one).
    const T* __restrict__ src = a.sortedP + (int64_t)min(d, DIM - 1) * a.mStride + off;
    if (a.tolerance) {
        const double wide = WideSum<T>::value(group_parallel_sum<T, kRow>(src, cnt));
        if (seg < 0 || t != 0) return;
        shift_coordinate_wide(a, d, slot, seg, cnt, wide);
        return;
    }
    const int steps = wave_row_max(cnt);
    const int rowInWave = (threadIdx.x & (kWave - 1)) / kRow;
    const T sum = group_ordered_sum<T, kRow>(src, cnt, steps, waveLds + rowInWave * (kRowTile + WideOf<T>::n), T(0));
    if (seg < 0 || t != 0) return;
    shift_coordinate(a, d, slot, seg, cnt, sum);
}

// Ordered integer of a value (the order of the values, -0 and +0 both 0) and back (never -0).
__device__ __forceinline__ int64_t ord_of(float x) {
    const int i = __float_as_int(x);
    return i >= 0 ? (int64_t)i : -(int64_t)(i & 0x7fffffff);
}
__device__ __forceinline__ int64_t ord_of(double x) {
    const long long i = __double_as_longlong(x);
    return i >= 0 ? (int64_t)i : -(int64_t)(i & 0x7fffffffffffffffLL);
}
template <typename T>
__device__ __forceinline__ T from_ord(int64_t o);
template <>
__device__ __forceinline__ float from_ord<float>(int64_t o) {
    return __int_as_float(o >= 0 ? (int)o : (int)(0x80000000u | (unsigned)(-o)));
}
template <>
__device__ __forceinline__ double from_ord<double>(int64_t o) {
    return __longlong_as_double(o >= 0 ? (long long)o : (long long)(0x8000000000000000ULL | (unsigned long long)(-o)));
}
__device__ __forceinline__ bool same_bits(float a, float b) { return __float_as_uint(a) == __float_as_uint(b); }
__device__ __forceinline__ bool same_bits(double a, double b) {
    return __double_as_longlong(a) == __double_as_longlong(b);
}

// Long chains: one block of kParts waves per (list entry, coordinate).  The serial sum of a
// chain is cut into kParts runs of consecutive members, one per wave, summed at the same time:
// wave p cannot know the exact partial sum its run starts from (the serial sum of every member
// before it), so each of its 64 lanes sums the run from its own start, the 64 consecutive values
// of the type centred on the correctly rounded sum of those members (from the wider sums of the
// runs before it).  Wave 0 then walks the runs in order: run p's result is the one of the lane
// whose start equals, bit for bit, the exact sum through run p - 1 -- the same adds of the same
// values in the same order as one serial sum, so the same result.  When no lane started from that
// value (the rounding errors of the serial sum before run p exceed 32 units of its last place,
// about one run in five hundred), wave 0 sums run p again from it.  The result never depends on
// the estimate; only the time does.
constexpr int kParts = 256 / kWave;

template <typename T>
struct PartLds {
    typename WideSum<T>::Acc wide[kParts];
    int64_t ordE[kParts];
    T cand[kParts][kWave];
    T start[kParts][kWave];
};

template <typename T, int DIM>
__device__ __forceinline__ void chain_sum_parts(const ShiftArgs<T>& a, int64_t entry, T* tile, PartLds<T>& sh) {
    using W = WideSum<T>;
    using Run = OrderedRun<T, kWave>;
    const int64_t e = entry / DIM;
    const int d = (int)(entry - e * DIM);
    // every exit before the barriers below is the same for the whole block.  The list entry is
    // loaded with the live count (in bounds either way: e < the list's capacity), and the chain's
    // range and segment together once the entry is known to be live.
    const int key = a.longList[e];
    if (a.longCount ? e >= min(*a.longCount, a.longCap) : (e << a.longShift) >= *a.mPtr) return;
    if (key < 0) return;
    const int slot = a.denseChains ? a.rowSlots[key] : key;
    const int seg = a.slotSeg[slot];
    const int off = a.chainOff[key];
    const int cnt = a.chainEnd[key] - off;
    // a chain listed by length belongs to a live segment and a live slot (the member scatter
    // lists no other); a chain listed by position is checked
    if (!a.longCount && (seg < 0 || a.segDone[seg] || (int)(slot - a.segs[seg].sOff) >= a.segK[seg])) return;
    const T* __restrict__ base = a.sortedP + (int64_t)d * a.mStride + off;
    const int w = threadIdx.x / kWave, lane = lane_id();
    if (a.tolerance) {
        // a quarter of the chain per wave in the wider type, combined in a fixed order
        const int q0 = (int)((int64_t)cnt * w / kParts), q1 = (int)((int64_t)cnt * (w + 1) / kParts);
        const typename W::Acc part = group_parallel_sum<T, kWave>(base + (q1 > q0 ? q0 : 0), q1 - q0);
        if (lane == 0) sh.wide[w] = part;
        __syncthreads();
        if (w != 0) return;
        typename W::Acc total = W::zero();
        for (int q = 0; q < kParts; ++q) total = W::add(total, sh.wide[q]);
        if (lane == 0) shift_coordinate_wide(a, d, slot, seg, cnt, W::value(total));
        return;
    }
    auto run_of = [&](int p, int& n) {
        const int b0 = (int)((int64_t)cnt * p / kParts);
        n = (int)((int64_t)cnt * (p + 1) / kParts) - b0;
        return base + (n > 0 ? b0 : 0);   // an empty run reads (and discards) the chain's first member
    };
    int n;
    const T* __restrict__ src = run_of(w, n);
    // the run's first tile, then the wider sum of the whole run from those registers and the
    // loads of the rest, issued before any of them is waited for
    Run run(src, n);
    run.load(0);
    typename W::Acc acc[4] = {W::zero(), W::zero(), W::zero(), W::zero()};
    for (int i0 = Run::kTile; i0 < n; i0 += Run::kTile) {
        T v[kLongPer];
#pragma unroll
        for (int c = 0; c < kLongPer; ++c) {
            const int i = i0 + c * kWave + lane;
            v[c] = i < n ? src[i] : T(0);
        }
#pragma unroll
        for (int c = 0; c < kLongPer; ++c) acc[c & 3] = W::add(acc[c & 3], v[c]);
    }
#pragma unroll
    for (int c = 0; c < kLongPer; ++c) acc[c & 3] = W::add(acc[c & 3], run.reg[c]);
    typename W::Acc wide = W::add(W::add(acc[0], acc[1]), W::add(acc[2], acc[3]));
#pragma unroll
    for (int m = kWave / 2; m >= 1; m /= 2) wide = W::add(wide, W::xorShfl(wide, m));
    if (lane == 0) sh.wide[w] = wide;
    __syncthreads();
    T start = T(0);
    if (w > 0) {
        typename W::Acc est = W::zero();
        for (int q = 0; q < w; ++q) est = W::add(est, sh.wide[q]);
        const T e0 = (T)W::value(est);
        const int64_t oe = ord_of(e0);
        start = isfinite(e0) ? from_ord<T>(oe + lane - kWave / 2) : T(0);
        if (lane == 0) sh.ordE[w] = oe;
    }
    const T cand = run.run(n, tile, start);
    sh.cand[w][lane] = cand;
    sh.start[w][lane] = start;
    __syncthreads();
    if (w != 0) return;
    T x = sh.cand[0][0];
    for (int p = 1; p < kParts; ++p) {
        const int64_t o = ord_of(x) - sh.ordE[p] + kWave / 2;
        if (o >= 0 && o < kWave && same_bits(sh.start[p][o], x)) {
            x = sh.cand[p][o];
        } else {
            int np;
            const T* __restrict__ sp = run_of(p, np);
            x = group_ordered_sum<T, kWave>(sp, np, np, tile, x);
        }
    }
    if (lane == 0) shift_coordinate(a, d, slot, seg, cnt, x);
}

// blocks [0, rowBlocks): 16-lane rows; the rest: one listed chain and coordinate each.  The
// convergence bookkeeping follows in a launch of its own (evolve_finalize_kernel).
template <typename T, int DIM>
__global__ __launch_bounds__(256) void chain_sum_shift_kernel(ShiftArgs<T> a, int64_t rowBlocks) {
    __shared__ __align__(16) T tiles[256 / kWave][wave_lds_elems<T>()];
    __shared__ PartLds<T> parts;
    if (*a.allDone) return;
    if ((int64_t)blockIdx.x < rowBlocks)
        chain_sum_rows<T, DIM>(a, blockIdx.x, tiles[threadIdx.x / kWave]);
    else
        chain_sum_parts<T, DIM>(a, (int64_t)blockIdx.x - rowBlocks, tiles[threadIdx.x / kWave], parts);
}

// ---------------------------------------------------------------------------------
// 5. convergence bookkeeping (one block)
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void evolve_finalize_kernel(int numSegs, int* __restrict__ segDone,
                                                               int* __restrict__ segMoved,
                                                               int* __restrict__ segIter, int maxIter,
                                                               int* __restrict__ allDone,
                                                               int* __restrict__ totalIters) {
    if (*allDone) return;
    finalize_segments(numSegs, segDone, segMoved, segIter, maxIter, allDone, totalIters);
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
template <typename T>
hipError_t launch_chain_sum_shift(const ShiftArgs<T>& a, int dim, int64_t maxM, bool gather, hipStream_t st) {
    const int64_t rowBlocks = ((int64_t)dim * a.totalSlots + 15) / 16;
    if (rowBlocks <= 0) return hipSuccess;
    const int64_t waves = a.longCount ? (int64_t)dim * a.longCap
                                      : a.longList ? (int64_t)dim * long_list_entries(maxM, a.longShift) : 0;
    const int64_t blocks = rowBlocks + waves;
    const unsigned gblocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((maxM + 255) / 256, 8192));
    bool ok = false;
    RD_DISPATCH_DIM(dim, {
        if (gather)
            hipLaunchKernelGGL((gather_members_kernel<T, DIM>), dim3(gblocks), dim3(256), 0, st, a.P, a.pStride,
                               a.pIndex, a.members, a.mPtr, a.sortedP, a.mStride, a.allDone);
        hipLaunchKernelGGL((chain_sum_shift_kernel<T, DIM>), dim3((unsigned)blocks), dim3(256), 0, st, a,
                           rowBlocks);
        ok = true;
    });
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

hipError_t launch_chain_bounds(const int* keys, const int64_t* dM, int64_t maxM, int* chainOff,
                               int* chainEnd, int* longList, int longShift, const int* allDone,
                               hipStream_t st) {
    int64_t blocks = std::min<int64_t>((maxM + 255) / 256, 8192);
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(chain_bounds_kernel, dim3((unsigned)blocks), dim3(256), 0, st, keys, dM, chainOff,
                       chainEnd, longList, longShift, allDone);
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

template hipError_t launch_chain_sum_shift<float>(const ShiftArgs<float>&, int, int64_t, bool, hipStream_t);
template hipError_t launch_chain_sum_shift<double>(const ShiftArgs<double>&, int, int64_t, bool, hipStream_t);
template hipError_t launch_fill_sentinel<float>(float*, int64_t, int, const SegDesc*, const int*, int, hipStream_t);
template hipError_t launch_fill_sentinel<double>(double*, int64_t, int, const SegDesc*, const int*, int, hipStream_t);

}  // namespace rdhip
