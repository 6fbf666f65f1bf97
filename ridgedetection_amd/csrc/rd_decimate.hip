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
//   4. scan     -- groups below kDecRoundsMin elements: one wave per group walks the
//                  elements in order, 64 at a time: the window's counters and CSR rows are
//                  staged in LDS, decisions are uniform, and a removal decrements its
//                  neighbours' counters (LDS for the window, device-scope atomics beyond it);
//      rounds   -- larger groups, one pass at a time: element e's test depends only on
//                  which of its earlier (lower-index) neighbours the pass removed, since
//                  its later neighbours are still live when the walk reaches it.  So in
//                  every round each undecided element resumes the scan of its earlier
//                  neighbours (rows are in increasing index order) and decides once all of
//                  them have decided; the earliest undecided element always can, so every
//                  round makes progress, and random-like orders need few rounds.  The
//                  group-wide count rule (stop below 3 / keep below 3) can only act when the
//                  pass would leave fewer than 3 points: such a pass is discarded and the
//                  group resumes in the serial walk from the pass-start state.  Otherwise
//                  the removals apply and the counts are recounted from the CSR rows;
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

// Bounding box of every chunk of kDecChunk consecutive elements: lo at [c][0, DIM), hi at
// [c][DIM, 2 DIM).  kBoxLanes lanes per chunk, each folding every kBoxLanes-th element.  The
// kernel also stages the coordinates in element order (E[e][0, DIM)) for the count.
constexpr int kDecChunk = 8;
constexpr int kBoxLanes = kDecChunk < kWave ? kDecChunk : kWave;
constexpr int kBoxChunksPerBlock = 256 / kBoxLanes;
static_assert((kDecChunk & (kDecChunk - 1)) == 0 && kDecChunk >= 8, "kDecChunk: a power of two");

template <typename T, int DIM>
__global__ __launch_bounds__(256) void dec_chunk_box_kernel(DecArgs a, const T* __restrict__ S, int64_t ss,
                                                            T* __restrict__ boxes, int64_t maxChunks,
                                                            T* __restrict__ E) {
    const int64_t c = (int64_t)blockIdx.x * kBoxChunksPerBlock + threadIdx.x / kBoxLanes;
    if (c >= maxChunks) return;   // whole chunk groups leave together: the shuffles stay within one
    const int64_t total = *a.totalElems;
    const int sub = threadIdx.x % kBoxLanes;
    T lo[DIM], hi[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        lo[d] = FpTraits<T>::big;
        hi[d] = -FpTraits<T>::big;
    }
    for (int k = sub; k < kDecChunk; k += kBoxLanes) {
        const int64_t e = c * kDecChunk + k;
        if (e >= total) break;
        const int se = a.elemSlot[e];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const T x = S[d * ss + se];
            E[e * DIM + d] = x;
            lo[d] = min(lo[d], x);
            hi[d] = max(hi[d], x);
        }
    }
#pragma unroll
    for (int m = 1; m < kBoxLanes; m <<= 1)
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            lo[d] = min(lo[d], __shfl_xor(lo[d], m));
            hi[d] = max(hi[d], __shfl_xor(hi[d], m));
        }
    if (sub == 0)
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            boxes[c * 2 * DIM + d] = lo[d];
            boxes[c * 2 * DIM + DIM + d] = hi[d];
        }
}

// Boxes of kDecSuper consecutive chunks (the hull of their boxes), after the chunk boxes.
constexpr int kDecSuper = 32;

template <typename T, int DIM>
__global__ __launch_bounds__(256) void dec_super_box_kernel(T* __restrict__ boxes, int64_t maxChunks,
                                                            int64_t maxSupers) {
    const int64_t sc = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (sc >= maxSupers) return;
    T lo[DIM], hi[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        lo[d] = FpTraits<T>::big;
        hi[d] = -FpTraits<T>::big;
    }
    for (int64_t c = sc * kDecSuper, c1 = min(maxChunks, (sc + 1) * kDecSuper); c < c1; ++c)
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            lo[d] = min(lo[d], boxes[c * 2 * DIM + d]);
            hi[d] = max(hi[d], boxes[c * 2 * DIM + DIM + d]);
        }
    T* out = boxes + (maxChunks + sc) * 2 * DIM;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
        out[d] = lo[d];
        out[DIM + d] = hi[d];
    }
}

// Row entries an element may keep from the count pass (provisional rows).
constexpr int kDecProv = 32;

// Neighbour counts within r2 and 2 r2 of every element in its group (fill = 0), then the
// CSR rows (fill = 1).  With prov, the count pass also writes each element's first kDecProv
// row entries to prov[e][.]; rows that fit are copied into place (dec_prov_copy_kernel) and
// the fill pass walks only the elements with longer rows.  A wave walks the union of its lanes' chunk ranges with wave-uniform
// indices, so candidate coordinates and boxes are broadcast loads; each lane tests only the
// candidates of its own group, in increasing element order, and only in chunks whose box
// passes its own test, as a walk of its own would.
template <typename T, int DIM>
__global__ __launch_bounds__(256) void dec_count_kernel(DecArgs a, const T* __restrict__ E,
                                                        const T* __restrict__ groupRSqr, int fill,
                                                        int64_t maxElems, const T* __restrict__ boxes,
                                                        int64_t maxChunks, int* __restrict__ prov) {
    const int64_t total = *a.totalElems;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= total && !fill && e <= maxElems) a.rowOff[e] = 0;   // tail of the offsets scan
    // the fill pass skips the rows the provisional copy already wrote
    const bool valid = e < total && !(fill && prov && a.c2[e] - 1 <= kDecProv);
    if (__ballot(valid) == 0) return;   // no element of the wave needs a walk
    int o = 0, n = 0;
    T rSqr = T(0), p[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) p[d] = T(0);
    if (valid) {
        const int g = a.elemGroup[e];
        o = a.groupOff[g];
        n = a.groupCnt[g];
        rSqr = groupRSqr[g];
#pragma unroll
        for (int d = 0; d < DIM; ++d) p[d] = E[e * DIM + d];
    }
    const T r4 = T(4) * rSqr;
    int c1 = 0, c2 = 0;
    int w = fill && valid ? a.rowOff[e] : 0;
    int np = 0;   // row entries found (count pass)
    // Chunks whose box lies farther than 2 r2 hold no neighbour: the squared distance to a
    // box bounds every member's from below; the margin covers the rounding of both sums.  A
    // super box holds its chunks' boxes, and every step of the box distance is monotone, so
    // a skipped super box skips only chunks that would be skipped one by one.
    const T skipAbove = r4 * (T(1) + T(16 * (DIM + 2)) * FpTraits<T>::unit_roundoff) + FpTraits<T>::tiny;
    auto boxDist = [&](const T* b) {
        T bd = T(0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const T t = max(max(b[d] - p[d], p[d] - b[DIM + d]), T(0));
            bd = bd + t * t;
        }
        return bd;
    };
    const bool any = valid && n > 0;
    const int cFirst = any ? o / kDecChunk : INT_MAX, cLast = any ? (o + n - 1) / kDecChunk : -1;
    int wFirst = cFirst, wLast = cLast;
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
        wFirst = min(wFirst, __shfl_xor(wFirst, m));
        wLast = max(wLast, __shfl_xor(wLast, m));
    }
    wFirst = __builtin_amdgcn_readfirstlane(wFirst);
    wLast = __builtin_amdgcn_readfirstlane(wLast);
    for (int sc = wFirst / kDecSuper; wFirst <= wLast && sc <= wLast / kDecSuper; ++sc) {
        const bool wantS = any && sc >= cFirst / kDecSuper && sc <= cLast / kDecSuper &&
                           !(boxDist(boxes + (maxChunks + sc) * 2 * DIM) > skipAbove);
        if (__ballot(wantS) == 0) continue;
        const int cEnd = min(wLast, sc * kDecSuper + kDecSuper - 1);
        for (int c = max(wFirst, sc * kDecSuper); c <= cEnd; ++c) {
            const bool wantC = wantS && c >= cFirst && c <= cLast && !(boxDist(boxes + (int64_t)c * 2 * DIM) > skipAbove);
            if (__ballot(wantC) == 0) continue;
            // E holds maxChunks * kDecChunk rows: a candidate past the end is read but never used
#pragma unroll 4
            for (int k = 0; k < kDecChunk; ++k) {
                const int j = c * kDecChunk + k;
                T q[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) q[d] = E[(int64_t)j * DIM + d];
                if (wantC && j >= o && j < o + n) {
                    const T dd = exact_sqdist<T, DIM>(p, q);
                    const bool in1 = dd <= rSqr;
                    const bool in2 = dd <= r4;
                    c1 += in1;
                    c2 += in2;
                    if (in2 && j != e) {
                        const int v = (j << 1) | (in1 ? 1 : 0);
                        if (fill)
                            a.nbr[w++] = v;
                        else if (prov && np < kDecProv)
                            prov[e * kDecProv + np] = v;
                        ++np;
                    }
                }
            }
        }
    }
    if (!fill && valid) {
        a.c1[e] = c1;
        a.c2[e] = c2;
        // neighbours other than itself (scanned to offsets); elements of a skipped group
        // (n == 0) have no list at all
        a.rowOff[e] = n > 0 ? c2 - 1 : 0;
    }
}

// Copies the provisional rows that hold the whole row (c2 - 1 <= kDecProv) into the CSR.
__global__ __launch_bounds__(256) void dec_prov_copy_kernel(DecArgs a, const int* __restrict__ prov) {
    const int64_t total = *a.totalElems;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const int g = a.elemGroup[e];
    if (a.groupCnt[g] <= 0) return;   // skipped group: no row
    const int cnt = a.c2[e] - 1;
    if (cnt > kDecProv) return;
    const int o = a.rowOff[e];
    for (int k = 0; k < cnt; ++k) a.nbr[o + k] = prov[e * kDecProv + k];
}

__global__ __launch_bounds__(64) void dec_scan_kernel(DecArgs a) {
    __shared__ int wc1[kWave], wc2[kWave], wal[kWave];
    __shared__ int rows[kRowCap];
    const int g = blockIdx.x;
    const int mode = a.groupMode[g];
    if (mode != kDecSerial && mode != kDecResume) return;
    const int o = a.groupOff[g], n = a.groupCnt[g];
    const int lane = lane_id();
    if (mode == kDecSerial)
        for (int e = o + lane; e < o + n; e += kWave) st_rel(&a.alive[e], 1);
    __threadfence();
    // a resumed group continues from the start of its current pass
    int cnt = mode == kDecSerial ? n : a.groupAlive[g];
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
    if (lane == 0) {
        a.groupLeft[g] = cnt;
        a.groupMode[g] = kDecDone;
    }
}

// ---- dependency rounds --------------------------------------------------------------
__global__ void dec_modes_kernel(DecArgs a) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= a.numGroups) return;
    const int n = a.groupCnt[g];
    const bool rounds = n >= kDecRoundsMin;
    a.groupMode[g] = rounds ? kDecRounds : kDecSerial;
    a.groupAlive[g] = n;
    a.groupRemoved[g] = 0;
    a.groupApply[g] = 0;
    a.groupLeft[g] = n;
    if (rounds) atomicAdd(&a.parFlags[1], 1);
}

// every element of a group in rounds mode: live, and undecided for the coming pass
__global__ void dec_pass_reset_kernel(DecArgs a, int first) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= *a.totalElems) return;
    if (a.groupMode[a.elemGroup[e]] != kDecRounds) return;
    if (first) a.alive[e] = 1;
    a.state[e] = 0;
    a.prog[e] = a.rowOff[e];
    a.rc1[e] = 0;
    a.rc2[e] = 0;
}

// One round: an undecided element resumes the scan of its earlier neighbours.  Those
// still undecided bound its counts: with r1 / r2 removals known and u1 / u2 undecided (r2
// and 2 r2 neighbourhoods), x1 lies in [c1 - r1 - u1, c1 - r1] and x2 in [c2 - r2 - u2,
// c2 - r2], so the test is already settled when it holds (or fails) at both ends -- in a
// dense neighbourhood every point with at least three later neighbours within r2 is
// removed whatever the earlier ones do.  An unsettled element looks again, at most
// kDecRetries times (neighbours decided meanwhile by other waves are then seen), before
// leaving it to the next round.  The decided prefix of the row is folded into prog / rc.
constexpr int kDecRetries = 48;

__device__ __forceinline__ bool dec_removes(int x1, int x2) { return x1 >= 4 || x2 < 3; }

__global__ __launch_bounds__(256) void dec_round_kernel(DecArgs a) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= *a.totalElems) return;
    if (a.groupMode[a.elemGroup[e]] != kDecRounds) return;
    if (!a.alive[e] || a.state[e] != 0) return;
    int q0 = a.prog[e];
    const int qe = a.rowOff[e + 1];
    int r1 = a.rc1[e], r2 = a.rc2[e];   // removals among the folded prefix [rowOff, q0)
    const int c1 = a.c1[e], c2 = a.c2[e];
    for (int attempt = 0;; ++attempt) {
        int k1 = 0, k2 = 0, u1 = 0, u2 = 0;   // removals / undecided past the prefix
        bool prefix = true;
        for (int q = q0; q < qe; ++q) {
            const int v = a.nbr[q];
            const int j = v >> 1;
            if (j > e) break;                 // rows are ascending: no earlier neighbours left
            int sj = 1;
            if (a.alive[j]) sj = ld_acq(&a.state[j]);
            if (sj == 0) {
                prefix = false;
                ++u2;
                u1 += v & 1;
            } else {
                const int rm = sj == 2 ? 1 : 0;
                if (prefix) {
                    q0 = q + 1;
                    r2 += rm;
                    r1 += rm & v;
                } else {
                    k2 += rm;
                    k1 += rm & v;
                }
            }
        }
        const int hi1 = c1 - r1 - k1, hi2 = c2 - r2 - k2;   // no further removals
        const int lo1 = hi1 - u1, lo2 = hi2 - u2;            // every undecided one removed
        // the test rises with x1 and falls with x2: it holds over the whole box iff it holds
        // at (lo1, hi2), and fails over it iff it fails at (hi1, lo2)
        const bool surelyRemoved = dec_removes(lo1, hi2);
        const bool surelyKept = !dec_removes(hi1, lo2);
        if (surelyRemoved || surelyKept) {
            st_rel(&a.state[e], surelyRemoved ? 2 : 1);
            return;
        }
        if (attempt + 1 == kDecRetries) {
            a.prog[e] = q0;
            a.rc1[e] = r1;
            a.rc2[e] = r2;
            st_rel(&a.parFlags[0], 1);
            return;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

__global__ __launch_bounds__(256) void dec_pass_sum_kernel(DecArgs a) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int g = -1;
    bool rm = false;
    if (e < *a.totalElems) {
        g = a.elemGroup[e];
        rm = a.groupMode[g] == kDecRounds && a.alive[e] && a.state[e] == 2;
    }
    // one atomic per wave when its elements share a group (the common case)
    const int g0 = __shfl(g, 0);
    const bool uniform = __all(g == g0);
    if (uniform) {
        const int c = __popcll(__ballot(rm));
        if (lane_id() == 0 && c > 0) atomicAdd(&a.groupRemoved[g0], c);
    } else if (rm) {
        atomicAdd(&a.groupRemoved[g], 1);
    }
}

__global__ void dec_pass_decide_kernel(DecArgs a) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= a.numGroups) return;
    a.groupApply[g] = 0;
    if (a.groupMode[g] != kDecRounds) return;
    const int n = a.groupAlive[g], R = a.groupRemoved[g];
    a.groupRemoved[g] = 0;
    if (n - R < 3) {                   // the count rule would act: redo this pass serially
        a.groupMode[g] = kDecResume;
        atomicAdd(&a.parFlags[2], 1);
        return;
    }
    a.groupApply[g] = 1;
    a.groupAlive[g] = n - R;
    a.groupLeft[g] = n - R;
    if (R > 0 && n - R > 3) {
        atomicAdd(&a.parFlags[1], 1);
    } else {
        a.groupMode[g] = kDecDone;
    }
}

__global__ void dec_pass_apply_kernel(DecArgs a) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= *a.totalElems) return;
    if (!a.groupApply[a.elemGroup[e]]) return;
    if (a.alive[e] && a.state[e] == 2) a.alive[e] = 0;
}

// counts for the next pass over the live elements (itself included)
__global__ void dec_recount_kernel(DecArgs a) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= *a.totalElems) return;
    if (a.groupMode[a.elemGroup[e]] != kDecRounds || !a.alive[e]) return;
    int c1 = 1, c2 = 1;
    for (int q = a.rowOff[e], qe = a.rowOff[e + 1]; q < qe; ++q) {
        const int v = a.nbr[q];
        if (a.alive[v >> 1]) {
            ++c2;
            c1 += v & 1;
        }
    }
    a.c1[e] = c1;
    a.c2[e] = c2;
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

namespace {
hipError_t dec_ordered_removal(DecArgs& a, int64_t maxElems, DecWorkspace& ws, hipStream_t st) {
    hipError_t err;
    const int64_t need = 5 * std::max<int64_t>(maxElems, 1) + 4 * (int64_t)a.numGroups + 4;
    if (need > ws.parCap) {
        if (ws.par) (void)hipFree(ws.par);
        ws.par = nullptr;
        ws.parCap = 0;
        if ((err = deviceMalloc((void**)&ws.par, sizeof(int) * (need + need / 4))) != hipSuccess) return err;
        ws.parCap = need + need / 4;
    }
    int* p = ws.par;
    const int64_t me = std::max<int64_t>(maxElems, 1);
    a.state = p; p += me;
    a.prog = p; p += me;
    a.rc1 = p; p += me;
    a.rc2 = p; p += me;
    p += me;
    a.groupMode = p; p += a.numGroups;
    a.groupAlive = p; p += a.numGroups;
    a.groupRemoved = p; p += a.numGroups;
    a.groupApply = p; p += a.numGroups;
    a.parFlags = p;
    const unsigned gb = (unsigned)((a.numGroups + 255) / 256);
    const unsigned eb = (unsigned)((me + 255) / 256);
    int flags[3] = {0, 0, 0};
    if ((err = hipMemsetAsync(a.parFlags, 0, sizeof(int) * 3, st)) != hipSuccess) return err;
    hipLaunchKernelGGL(dec_modes_kernel, dim3(gb), dim3(256), 0, st, a);
    hipLaunchKernelGGL(dec_scan_kernel, dim3(a.numGroups), dim3(kWave), 0, st, a);
    if ((err = hipMemcpyAsync(flags, a.parFlags, sizeof(int) * 3, hipMemcpyDeviceToHost, st)) != hipSuccess) return err;
    if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
    bool first = true;
    while (flags[1] > 0) {
        if ((err = hipMemsetAsync(a.parFlags, 0, sizeof(int) * 3, st)) != hipSuccess) return err;
        hipLaunchKernelGGL(dec_pass_reset_kernel, dim3(eb), dim3(256), 0, st, a, first ? 1 : 0);
        first = false;
        // rounds until no element is left undecided; every round decides at least the
        // earliest undecided element of each group
        for (int64_t done = 0; done <= maxElems; done += 8) {
            for (int r = 0; r < 7; ++r) hipLaunchKernelGGL(dec_round_kernel, dim3(eb), dim3(256), 0, st, a);
            if ((err = hipMemsetAsync(a.parFlags, 0, sizeof(int), st)) != hipSuccess) return err;
            hipLaunchKernelGGL(dec_round_kernel, dim3(eb), dim3(256), 0, st, a);
            if ((err = hipMemcpyAsync(flags, a.parFlags, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess)
                return err;
            if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
            if (!flags[0]) break;
        }
        if (flags[0]) return hipErrorUnknown;   // unreachable: rounds always make progress
        hipLaunchKernelGGL(dec_pass_sum_kernel, dim3(eb), dim3(256), 0, st, a);
        hipLaunchKernelGGL(dec_pass_decide_kernel, dim3(gb), dim3(256), 0, st, a);
        hipLaunchKernelGGL(dec_pass_apply_kernel, dim3(eb), dim3(256), 0, st, a);
        hipLaunchKernelGGL(dec_recount_kernel, dim3(eb), dim3(256), 0, st, a);
        if ((err = hipMemcpyAsync(flags, a.parFlags, sizeof(int) * 3, hipMemcpyDeviceToHost, st)) != hipSuccess)
            return err;
        if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
        if (flags[2] > 0) hipLaunchKernelGGL(dec_scan_kernel, dim3(a.numGroups), dim3(kWave), 0, st, a);
    }
    return hipGetLastError();
}
}  // namespace

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
    // 2. chunk boxes, counts
    const int64_t maxChunks = std::max<int64_t>((maxElems + kDecChunk - 1) / kDecChunk, 1);
    const int64_t maxSupers = (maxChunks + kDecSuper - 1) / kDecSuper;
    // chunk boxes, super boxes, the element-order coordinates (maxChunks * kDecChunk rows),
    // then the provisional rows when they fit in 256 MB
    const int64_t coordBytes = ((maxChunks + maxSupers) * 2 * dim + maxChunks * kDecChunk * dim) * (int64_t)sizeof(T);
    const int64_t provOff = (coordBytes + 255) / 256 * 256;
    const int64_t provBytes = maxElems * kDecProv * (int64_t)sizeof(int);
    const bool useProv = provBytes <= (int64_t(256) << 20);
    const int64_t boxBytes = useProv ? provOff + provBytes : coordBytes;
    if (boxBytes > ws.boxCap) {
        if (ws.boxes) (void)hipFree(ws.boxes);
        ws.boxes = nullptr;
        ws.boxCap = 0;
        if ((err = deviceMalloc((void**)&ws.boxes, boxBytes + boxBytes / 4)) != hipSuccess) return err;
        ws.boxCap = boxBytes + boxBytes / 4;
    }
    T* boxes = static_cast<T*>(ws.boxes);
    T* coords = boxes + (maxChunks + maxSupers) * 2 * dim;
    int* prov = useProv ? reinterpret_cast<int*>(static_cast<char*>(ws.boxes) + provOff) : nullptr;
    const unsigned blocks = (unsigned)((maxElems + 1 + 255) / 256);
    bool ok = false;
    RD_DISPATCH_DIM(dim, {
        hipLaunchKernelGGL((dec_chunk_box_kernel<T, DIM>),
                           dim3((unsigned)((maxChunks + kBoxChunksPerBlock - 1) / kBoxChunksPerBlock)), dim3(256), 0, st, a,
                           (const T*)S, sStride, boxes, maxChunks, coords);
        hipLaunchKernelGGL((dec_super_box_kernel<T, DIM>), dim3((unsigned)((maxSupers + 255) / 256)), dim3(256), 0,
                           st, boxes, maxChunks, maxSupers);
        hipLaunchKernelGGL((dec_count_kernel<T, DIM>), dim3(blocks), dim3(256), 0, st, a, (const T*)coords,
                           groupRSqr, 0, maxElems, (const T*)boxes, maxChunks, prov);
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
        ws.nbr = nullptr;
        ws.nbrCap = 0;
        if ((err = deviceMalloc((void**)&ws.nbr, sizeof(int) * (nbrTotal + nbrTotal / 4 + 1024))) != hipSuccess)
            return err;
        ws.nbrCap = nbrTotal + nbrTotal / 4 + 1024;
    }
    a.nbr = ws.nbr;
    if (prov)
        hipLaunchKernelGGL(dec_prov_copy_kernel, dim3((unsigned)((maxElems + 255) / 256)), dim3(256), 0, st, a,
                           (const int*)prov);
    RD_DISPATCH_DIM(dim, {
        hipLaunchKernelGGL((dec_count_kernel<T, DIM>), dim3(blocks), dim3(256), 0, st, a, (const T*)coords,
                           groupRSqr, 1, maxElems, (const T*)boxes, maxChunks, prov);
    });
    // 4. ordered removal: small groups one wave each; large groups in dependency rounds
    if ((err = dec_ordered_removal(a, maxElems, ws, st)) != hipSuccess) return err;
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
