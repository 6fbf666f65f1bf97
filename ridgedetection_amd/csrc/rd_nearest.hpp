// rd_nearest.hpp -- nearest chosen point of a lane's samples (the evolve assignment of
// rd/cpu/brute_force/evolve.hpp:32-77), shared by the batch-wide scan (rd_evolve.hip, chosen
// points in HBM read with uniform scalar loads) and the segment-local iteration
// (rd_local.hip, chosen points in LDS).  Three screens, each followed by an exact decision
// that returns the reference's label: the direct one (nearest_labels, below), the expanded
// one (nearest_labels_x) and the expanded one on the matrix cores (nearest_labels_mfma).
//
// Direct screen: a screening distance (1 mul + DIM-1 fma) with a v_min3 tree tracks, per sample, the best
// tile-minimum over tiles of kChosenTile chosen points, the tile holding it and the second
// best tile-minimum.  The epilogue re-scans the winning tile for the winner and the in-tile
// runner-up, then decides exactly: the screening value differs from the reference
// distance by at most delta = 4(DIM+1)u relative, so if the runner-up exceeds
// best(1+3 delta) the screened winner is the reference argmin and only its exact distance
// is computed (for the r-test); otherwise (near ties) the lane falls back to the exact
// reference scan with strict-< first-index ties.
#pragma once

#include "rd_common.hpp"

namespace rdhip {

template <typename T, int DIM>
__device__ __forceinline__ T screen_dist(const T (&p)[DIM], const T* s /* s[d] */) {
    T t0 = s[0] - p[0];
    T a = t0 * t0;
#pragma unroll
    for (int d = 1; d < DIM; ++d) {
        T t = s[d] - p[d];
        a = fma(t, t, a);
    }
    return a;
}

template <typename T>
__device__ __forceinline__ T tile_min(T m, T a, T b);
template <>
__device__ __forceinline__ float tile_min<float>(float m, float a, float b) { return min3f(m, a, b); }
template <>
__device__ __forceinline__ double tile_min<double>(double m, double a, double b) {
    return fmin(fmin(m, a), b);
}

// Expanded-form screening value of one pair (see nearest_labels_x below).
template <typename T, int DIM>
__device__ __forceinline__ T x_value(const T (&q2)[DIM], const T* s /* s[d] */, T b) {
    T v = fma(s[0], q2[0], b);
#pragma unroll
    for (int d = 1; d < DIM; ++d) v = fma(s[d], q2[d], v);
    return v;
}

// ---------------------------------------------------------------------------------
// Main loops: per sample the best tile-minimum over tiles of kChosenTile chosen points, the
// tile holding it (wt) and the second-best tile-minimum.  w holds the samples (direct screen)
// or -2(p - o) (expanded screen, X); Tb is the chosen rows (direct) or the expanded table.  The
// chosen values stream through SGPRs with uniform scalar loads, a whole tile per wait.
template <typename T, int PPT>
__device__ __forceinline__ void tile_done(int tj, const T (&m)[PPT], T (&best)[PPT], T (&second)[PPT], int (&wt)[PPT]) {
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const bool c = m[j] < best[j];
        wt[j] = c ? tj : wt[j];
        second[j] = second_update(best[j], second[j], m[j]);
        best[j] = dmin(best[j], m[j]);
    }
}

template <typename T, int DIM, int PPT, bool X>
__device__ __forceinline__ void screen_main(const T (&w)[PPT][DIM], int K, const T* __restrict__ Tb, int64_t ss,
                                            T (&best)[PPT], T (&second)[PPT], int (&wt)[PPT]) {
    constexpr T big = FpTraits<T>::big;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        best[j] = big;
        second[j] = big;
        wt[j] = 0;
    }
    const int nT = (K + kChosenTile - 1) / kChosenTile;
    if constexpr (X) {
        for (int tj = 0; tj < nT; ++tj) {
            T s[kChosenTile][DIM];
            T b[kChosenTile];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
#pragma unroll
                for (int t = 0; t < kChosenTile; ++t) s[t][d] = Tb[d * ss + tj * kChosenTile + t];
#pragma unroll
            for (int t = 0; t < kChosenTile; ++t) b[t] = Tb[DIM * ss + tj * kChosenTile + t];
            T m[PPT];
#pragma unroll
            for (int j = 0; j < PPT; ++j) m[j] = big;
#pragma unroll
            for (int t = 0; t < kChosenTile; t += 2) {
#pragma unroll
                for (int j = 0; j < PPT; ++j) {
                    const T a0 = x_value<T, DIM>(w[j], s[t], b[t]);
                    const T a1 = x_value<T, DIM>(w[j], s[t + 1], b[t + 1]);
                    m[j] = tile_min<T>(m[j], a0, a1);
                }
            }
            tile_done<T, PPT>(tj, m, best, second, wt);
        }
    } else {
        for (int tj = 0; tj < nT; ++tj) {
            T s[kChosenTile][DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
#pragma unroll
                for (int t = 0; t < kChosenTile; ++t) s[t][d] = Tb[d * ss + tj * kChosenTile + t];
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                T m = big;
#pragma unroll
                for (int t = 0; t < kChosenTile; t += 2) {
                    T a0 = screen_dist<T, DIM>(w[j], s[t]);
                    T a1 = screen_dist<T, DIM>(w[j], s[t + 1]);
                    m = tile_min<T>(m, a0, a1);
                }
                const bool c = m < best[j];
                wt[j] = c ? tj : wt[j];
                second[j] = second_update(best[j], second[j], m);
                best[j] = dmin(best[j], m);
            }
        }
    }
}

// label[j] = index k in [0, K) of the reference's nearest chosen point of sample p[j] when
// it lies within r (squared distance <= rSqr), else -1.  Chosen point k, coordinate d is
// Sb[d * ss + k]; slots K .. (tile end) hold the far-away sentinel.  Lanes with
// valid[j] == false get -1.  Every lane of the wave must call this (the near-tie fallback
// votes with a ballot).
template <typename T, int DIM, int PPT>
__device__ __forceinline__ void nearest_labels(const T (&p)[PPT][DIM], const bool (&valid)[PPT], int K,
                                               const T* __restrict__ Sb, int64_t ss, T rSqr, int (&label)[PPT]) {
#pragma unroll
    for (int j = 0; j < PPT; ++j) label[j] = -1;
    if (K <= 0) return;
    constexpr T big = FpTraits<T>::big;
    T best[PPT], second[PPT];
    int wt[PPT];
    screen_main<T, DIM, PPT, false>(p, K, Sb, ss, best, second, wt);

    // epilogue: winner inside the winning tile, runner-up, exact decision
    const T delta = T(4 * (DIM + 1)) * FpTraits<T>::unit_roundoff;
    const T tiny = FpTraits<T>::tiny;
    bool amb[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        amb[j] = false;
        if (!valid[j]) continue;
        if (!(best[j] <= rSqr * (T(1) + T(2) * delta) + tiny)) continue;   // certainly not a member
        int kwin = -1;
        T secIn = big;
        const int t0 = wt[j] * kChosenTile;
        for (int t = 0; t < kChosenTile; ++t) {
            T sv[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d) sv[d] = Sb[d * ss + t0 + t];
            T av = screen_dist<T, DIM>(p[j], sv);
            if (kwin < 0 && av == best[j])
                kwin = t;
            else
                secIn = dmin(secIn, av);
        }
        const T sec = dmin(second[j], secIn);
        if (kwin < 0 || !(sec > best[j] * (T(1) + T(3) * delta) + tiny)) {
            amb[j] = true;
            continue;
        }
        const int k = t0 + kwin;
        T sv[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) sv[d] = Sb[d * ss + k];
        T e = exact_sqdist<T, DIM>(p[j], sv);
        if (e <= rSqr) label[j] = k;
    }

    // exact reference scan for near ties (rare)
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        if (__ballot(amb[j]) == 0) continue;
        if (amb[j]) {
            T minSq = big;
            int minK = -1;
            for (int k = 0; k < K; ++k) {
                T sv[DIM];
#pragma unroll
                for (int d = 0; d < DIM; ++d) sv[d] = Sb[d * ss + k];
                T e = exact_sqdist<T, DIM>(p[j], sv);
                if (e < minSq) {
                    minSq = e;
                    minK = k;
                }
            }
            label[j] = (minK >= 0 && minSq <= rSqr) ? minK : -1;
        }
    }
}

// ---------------------------------------------------------------------------------
// Expanded-form screen.  With a centre o per segment, s' = s - o and q = p - o,
//   |s - p|^2 = |s'|^2 - 2 s'.q + |q|^2,
// and |q|^2 is the same for every chosen point of a sample, so the argmin of
//   V(c) = b_c + sum_d s'_cd (-2 q_d),   b_c = |s'_c|^2
// is the nearest chosen point: DIM fused multiply-adds per pair (b_c moves to a VGPR once per
// chosen point and serves the lane's PPT samples) against 2 DIM operations for the direct form.
// The table rows s'_c and b_c come from expand_chosen (rd_screen.hip): s'_c is s_c - o
// rounded once, b_c is |s'_c|^2 rounded once (FP32: summed in double; FP64: DIM roundings),
// and the sentinel slots past K hold s' = 0, b = FpTraits<T>::xSentinel.
//
// Error bound.  With u the unit roundoff, B = max_c b_c over the live chosen points and
// N = fl(|q|^2), every computed V(c) satisfies
//   |V(c) + |q|^2 - D(c)| <= E = 1.001 u (cb B + (DIM + 5) N),   cb = 2 DIM + 6 (FP32), 3 DIM + 6 (FP64),
// where D(c) = |s_c - p|^2 exactly: DIM u (2B + N) from the fused chain (its partial sums are
// at most b + 2 sum |s'_d||q_d| <= 2B + N), u B (FP64: DIM u B) from rounding b_c, and
// 4 u (B + N) from rounding s' and q (|s'_c - q| is within u(|s'_c| + |q|) of |s_c - p|).
// The reference distance D_ref(c) (rd_inner.hpp:41-49 order, exact_sqdist) is within
// gam = (DIM + 3) u of D(c) relatively.  Let c* be the screened winner, V1 = V(c*), V2 the
// smallest V of the other chosen points and e = D_ref(c*).  Then for every c != c*,
//   D(c) - D(c*) >= V2 - V1 - 2E,
// so V2 - V1 > 2E + 2.02 gam e makes c* the unique reference argmin (strictly smaller D_ref
// than every other chosen point); its label follows from e <= r^2.  Otherwise the reference
// argmin c0 has V(c0) <= V1 + 2E + 2.02 gam e, and an exact pass over the chosen points inside
// that window (ascending index, strict <) finds it with the reference's first-index rule.
// Samples whose lower bound (V1 + N - E')(1 - 2 gam) on every D_ref exceeds r^2 are outside
// every ball (E' = E + 1.01 u (4 B + (DIM + 5) N) adds the gap between N and |q|^2 and the
// rounding of that bound itself).

// The decision after a main loop (see the error bound above): winner and in-tile runner-up
// from the winning tile, the acceptance test, and the exact window pass for the near ties.
// MT: tiles of the matrix-core main loop (nearest_labels_mfma) instead of contiguous ones.
template <bool MT>
__device__ __forceinline__ int tile_slot(int tile, int t) {
    if constexpr (MT)
        return (tile >> 1) * 32 + 4 * (tile & 1) + (t & 3) + 8 * (t >> 2);
    else
        return tile * kChosenTile + t;
}

template <typename T, int DIM, int PPT, bool MT>
__device__ __forceinline__ void x_decide(const T (&p)[PPT][DIM], const T (&q2)[PPT][DIM], const T (&nq)[PPT],
                                         const bool (&valid)[PPT], int K, const T* __restrict__ Sb,
                                         const T* __restrict__ Xb, int64_t ss, T bmax, T rSqr, const T (&best)[PPT],
                                         const T (&second)[PPT], const int (&wt)[PPT], int (&label)[PPT]) {
    constexpr T big = FpTraits<T>::big;
    constexpr T u = FpTraits<T>::unit_roundoff;
    const int nT = (K + kChosenTile - 1) / kChosenTile;
    const int kEnd = nT * kChosenTile;
    constexpr T cb = sizeof(T) == 4 ? T(2 * DIM + 6) : T(3 * DIM + 6);
    const T gam = T(DIM + 3) * u;
    const T tiny = FpTraits<T>::tiny;
    bool amb[PPT];
    T window[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        amb[j] = false;
        window[j] = big;
        if (!valid[j]) continue;
        const T E = T(1.001) * u * (cb * bmax + T(DIM + 5) * nq[j]) + T(8 * (DIM + 1)) * FpTraits<T>::minNormal;
        const T Eout = E + T(1.01) * u * (T(4) * bmax + T(DIM + 5) * nq[j]);
        if ((best[j] + nq[j] - Eout) * (T(1) - T(2) * gam) > rSqr) continue;   // certainly not a member
        int kwin = -1;
        T secIn = big;
        // the winning tile as 4 runs of 4 consecutive slots (contiguous tiles: one run of 16),
        // so the loads of a run merge into wide ones; runs past the tile-rounded count (matrix-
        // core tiles reach them) hold the sentinel, and their loads read slots 0..3 instead so
        // that every load stays in bounds
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int base = tile_slot<MT>(wt[j], 4 * g);
            const bool live = base < kEnd;
            const int lb = live ? base : 0;
            T sv[4][DIM + 1];
#pragma unroll
            for (int d = 0; d <= DIM; ++d)
#pragma unroll
                for (int t = 0; t < 4; ++t) sv[t][d] = Xb[d * ss + lb + t];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const T xv = x_value<T, DIM>(q2[j], sv[t], sv[t][DIM]);
                const T av = live ? xv : FpTraits<T>::xSentinel;
                if (kwin < 0 && av == best[j])
                    kwin = 4 * g + t;
                else
                    secIn = dmin(secIn, av);
            }
        }
        const int k = kwin < 0 ? K : tile_slot<MT>(wt[j], kwin);
        if (k >= K) {
            amb[j] = true;   // not reached for finite data; the exact pass decides
            continue;
        }
        T sv[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) sv[d] = Sb[d * ss + k];
        const T e = exact_sqdist<T, DIM>(p[j], sv);
        const T thr = T(2) * E + T(2.02) * gam * e + tiny;
        const T sec = dmin(second[j], secIn);
        if (!(sec - best[j] > thr)) {
            amb[j] = true;
            // rounding of this sum is covered by the extra 0.01 thr + 4 u |V1|
            window[j] = best[j] + (T(1.01) * thr + T(4) * u * fabs(best[j]));
            continue;
        }
        if (e <= rSqr) label[j] = k;
    }

    // exact pass over the window of the near ties (see above)
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        if (__ballot(amb[j]) == 0) continue;
        if (amb[j]) {
            T minSq = big;
            int minK = -1;
            for (int tj = 0; tj < nT; ++tj) {
                T v[kChosenTile];
                T mt = big;
#pragma unroll
                for (int t = 0; t < kChosenTile; ++t) {
                    T sv[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) sv[d] = Xb[d * ss + tj * kChosenTile + t];
                    v[t] = x_value<T, DIM>(q2[j], sv, Xb[DIM * ss + tj * kChosenTile + t]);
                    mt = dmin(mt, v[t]);
                }
                if (mt <= window[j]) {
#pragma unroll
                    for (int t = 0; t < kChosenTile; ++t) {
                        const int k = tj * kChosenTile + t;
                        if (k < K && v[t] <= window[j]) {
                            T sv[DIM];
#pragma unroll
                            for (int d = 0; d < DIM; ++d) sv[d] = Sb[d * ss + k];
                            const T e = exact_sqdist<T, DIM>(p[j], sv);
                            if (e < minSq) {
                                minSq = e;
                                minK = k;
                            }
                        }
                    }
                }
            }
            label[j] = (minK >= 0 && minSq <= rSqr) ? minK : -1;
        }
    }
}


// label[j] as nearest_labels; Xb: the expanded table of the segment (row d < DIM: s'_d,
// row DIM: b), o: the segment's centre, bmax: B above.
template <typename T, int DIM, int PPT>
__device__ __forceinline__ void nearest_labels_x(const T (&p)[PPT][DIM], const bool (&valid)[PPT], int K,
                                                 const T* __restrict__ Sb, const T* __restrict__ Xb, int64_t ss,
                                                 const T (&o)[DIM], T bmax, T rSqr, int (&label)[PPT]) {
#pragma unroll
    for (int j = 0; j < PPT; ++j) label[j] = -1;
    if (K <= 0) return;
    constexpr T big = FpTraits<T>::big;
    constexpr T u = FpTraits<T>::unit_roundoff;
    T q2[PPT][DIM], nq[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        T n = T(0);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const T q = p[j][d] - o[d];
            q2[j][d] = T(-2) * q;
            n = fma(q, q, n);
        }
        nq[j] = n;
    }
    T best[PPT], second[PPT];
    int wt[PPT];
    screen_main<T, DIM, PPT, true>(q2, K, Xb, ss, best, second, wt);
    x_decide<T, DIM, PPT, false>(p, q2, nq, valid, K, Sb, Xb, ss, bmax, rSqr, best, second, wt, label);
}

// ---------------------------------------------------------------------------------
// Expanded-form screen on the matrix cores (FP32).  A wave screens CB column blocks of 32
// samples against blocks of 32 chosen points with v_mfma_f32_32x32x2f32, D = A B with
// A[i][k] the chosen point i's k-slot of [b, s'_0, ..., s'_{DIM-1}] (zero-padded to an even
// count) and B[k][j] sample j's k-slot of [1, -2q_0, ..., -2q_{DIM-1}].  The instruction
// forms each output as fma(a1, b1, fma(a0, b0, c)) over its two k-slots (checked on the
// device over millions of random operands), so the chain of ceil((DIM + 1) / 2) of them from
// c = 0 yields bit for bit x_value's fma(s'_{DIM-1}, -2q_{DIM-1}, ... fma(s'_0, -2q_0, b)):
// the decision (x_decide) and its bound are those of the VALU form.  (The bound's absolute
// term covers products the matrix cores might flush to zero below the smallest normal.)
// Operand layout: lane l supplies A[l % 32][l / 32] and B[l / 32][l % 32]; register r of lane l
// holds D[(r % 4) + 8 (r / 4) + 4 (l / 32)][l % 32].  Lane l's 16 values of a block are its
// column's sample against 16 chosen points: its tile (2 block + l / 32, tile_slot<true>).
typedef float rd_floatx16 __attribute__((ext_vector_type(16)));

template <int DIM>
constexpr int mfma_steps() { return (DIM + 2) / 2; }

// pc[c][d]: coordinate d of column block c's sample in column l % 32 (both half-waves hold it);
// label[e]: the label of the lane's e-th sample, column block 2 e + l / 32, column l % 32.
// Every lane of the wave must call this.
template <int DIM, int CB>
__device__ __forceinline__ void nearest_labels_mfma(const float (&pc)[CB][DIM], const bool (&validc)[CB], int K,
                                                    const float* __restrict__ Sb, const float* __restrict__ Xb,
                                                    int64_t ss, const float (&o)[DIM], float bmax, float rSqr,
                                                    int (&label)[CB / 2]) {
    static_assert(CB % 2 == 0, "column blocks come in pairs");
    constexpr int PE = CB / 2;
    constexpr int MK = mfma_steps<DIM>();
    constexpr float big = FpTraits<float>::big;
#pragma unroll
    for (int e = 0; e < PE; ++e) label[e] = -1;
    if (K <= 0) return;
    const int lane = threadIdx.x & 63, row = lane & 31, h = lane >> 5;
    float bq[CB][MK];
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
        for (int m = 0; m < MK; ++m) {
            const int ks = 2 * m + h;
            float v = 0.f;
            if (ks == 0) v = 1.f;
#pragma unroll
            for (int d = 0; d < DIM; ++d)
                if (ks == d + 1) v = -2.f * (pc[c][d] - o[d]);
            bq[c][m] = v;
        }
    const int kEnd = (K + kChosenTile - 1) / kChosenTile * kChosenTile;
    const int nB = (K + 31) / 32;
    float best[CB], second[CB];
    int wt[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        best[c] = big;
        second[c] = big;
        wt[c] = 0;
    }
    // A operands of a block: the lane's chosen point (row) in its k-slots; slots past the
    // tile-rounded count and the zero padding read table entry 0 (in bounds) and are replaced
    // by the sentinel / zero, without branches.  The next block's operands are loaded while
    // the current block is screened (vector loads complete in order).
    auto loadA = [&](int blk, float (&av)[MK]) {
        const int slot = blk * 32 + row;
#pragma unroll
        for (int m = 0; m < MK; ++m) {
            const int ks = 2 * m + h;
            const int r = ks == 0 ? DIM : ks - 1;   // table row of the k-slot
            const bool live = ks <= DIM && slot < kEnd;
            const float v = Xb[live ? r * ss + slot : 0];
            av[m] = live ? v : (ks == 0 ? FpTraits<float>::xSentinel : 0.f);
        }
    };
    float an[MK];
    loadA(0, an);
    for (int blk = 0; blk < nB; ++blk) {
        float av[MK];
#pragma unroll
        for (int m = 0; m < MK; ++m) av[m] = an[m];
        loadA(blk + 1 < nB ? blk + 1 : blk, an);
        const int tile = 2 * blk + h;
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            rd_floatx16 acc;
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
            for (int m = 0; m < MK; ++m) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[m], bq[c][m], acc, 0, 0, 0);
            float mn = fminf(acc[0], acc[1]);
#pragma unroll
            for (int r = 2; r < 16; r += 2) mn = min3f(mn, acc[r], acc[r + 1]);
            const bool w = mn < best[c];
            wt[c] = w ? tile : wt[c];
            second[c] = second_update(best[c], second[c], mn);
            best[c] = dmin(best[c], mn);
        }
    }
    // each lane takes one column block of every pair: the two half-waves' tiles of that column
    // block combine (smallest, second smallest and the tile of the smallest)
    float p[PE][DIM], q2[PE][DIM], nq[PE], B1[PE], S1[PE];
    int W1[PE];
    bool valid[PE];
#pragma unroll
    for (int e = 0; e < PE; ++e) {
        const float mb = h ? best[2 * e + 1] : best[2 * e];
        const float ms = h ? second[2 * e + 1] : second[2 * e];
        const int mw = h ? wt[2 * e + 1] : wt[2 * e];
        const float ob = __shfl_xor(h ? best[2 * e] : best[2 * e + 1], 32);
        const float os = __shfl_xor(h ? second[2 * e] : second[2 * e + 1], 32);
        const int ow = __shfl_xor(h ? wt[2 * e] : wt[2 * e + 1], 32);
        B1[e] = fminf(mb, ob);
        W1[e] = mb <= ob ? mw : ow;
        S1[e] = fminf(fmaxf(mb, ob), fminf(ms, os));
        valid[e] = h ? validc[2 * e + 1] : validc[2 * e];
        float n = 0.f;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            p[e][d] = h ? pc[2 * e + 1][d] : pc[2 * e][d];
            const float q = p[e][d] - o[d];
            q2[e][d] = -2.f * q;
            n = fma(q, q, n);
        }
        nq[e] = n;
    }
    x_decide<float, DIM, PE, true>(p, q2, nq, valid, K, Sb, Xb, ss, bmax, rSqr, B1, S1, W1, label);
}

}  // namespace rdhip
