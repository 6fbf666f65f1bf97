// rd_nearest.hpp -- nearest chosen point of a lane's samples (the evolve assignment of
// rd/cpu/brute_force/evolve.hpp:32-77), shared by the batch-wide scan (rd_evolve.hip, chosen
// points in HBM read with uniform scalar loads) and the segment-local iteration
// (rd_local.hip, chosen points in LDS).
//
// A screening distance (1 mul + DIM-1 fma) with a v_min3 tree tracks, per sample, the best
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
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        best[j] = big;
        second[j] = big;
        wt[j] = 0;
    }
    const int nT = (K + kChosenTile - 1) / kChosenTile;
    for (int tj = 0; tj < nT; ++tj) {
        T s[kChosenTile][DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
#pragma unroll
            for (int t = 0; t < kChosenTile; ++t) s[t][d] = Sb[d * ss + tj * kChosenTile + t];
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            T m = big;
#pragma unroll
            for (int t = 0; t < kChosenTile; t += 2) {
                T a0 = screen_dist<T, DIM>(p[j], s[t]);
                T a1 = screen_dist<T, DIM>(p[j], s[t + 1]);
                m = tile_min<T>(m, a0, a1);
            }
            const bool c = m < best[j];
            wt[j] = c ? tj : wt[j];
            second[j] = second_update(best[j], second[j], m);
            best[j] = dmin(best[j], m);
        }
    }

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

}  // namespace rdhip
