// ref_driver.cpp -- C ABI over the reference's own header-only CPU implementation.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the reference headers
// where they lie (`/root/reference/ridge_detection`), into oracle/_ref/.  It writes the golden
// fixtures under tests/golden/ (make_golden.py), which pin the restatement in rd_oracle.cpp
// (tests/test_oracle_golden.py), and -- in its OpenMP build -- it is the timed CPU baseline
// of bench.py.  Nothing in the product loads it.
//
// The reference copies coordinates rather than indices; indices are recovered here by
// matching coordinates back to the input (inputs used with this driver have no
// duplicate points).

#include "rd/cpu/brute_force/ridge_detection.hpp"
#include "rd/cpu/brute_force/choose.hpp"
#include "rd/cpu/brute_force/evolve.hpp"
#include "rd/cpu/brute_force/decimate.hpp"
#include "rd/cpu/brute_force/order.hpp"
#include "rd/utils/assessment_quality.hpp"
#include "rd/cpu/tiled/rd_tiled.hpp"
#include "rd/cpu/tiled/local_tile_tree.hpp"
#include "rd/cpu/tiled/grouped_tile_tree.hpp"

#include <cstring>
#include <list>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

template <typename T>
struct CoordIndex {
    std::unordered_map<std::string, long> m;
    int dim;
    CoordIndex(const T* P, long n, int d) : dim(d) {
        m.reserve((size_t)n * 2);
        for (long i = n - 1; i >= 0; --i)
            m[std::string(reinterpret_cast<const char*>(P + i * d), sizeof(T) * d)] = i;
    }
    long operator()(const T* p) const {
        auto it = m.find(std::string(reinterpret_cast<const char*>(p), sizeof(T) * dim));
        return it == m.end() ? -1 : it->second;
    }
};

template <typename T>
long refChoose(const T* P, long n, int dim, T r, T* S, long* idx) {
    std::vector<T> buf((size_t)n * dim);
    std::list<T*> lst;
    size_t cnt = 0;
    rd::choose(P, buf.data(), lst, (size_t)n, cnt, (size_t)dim, r);
    std::memcpy(S, buf.data(), sizeof(T) * cnt * dim);
    if (idx) {
        CoordIndex<T> ci(P, n, dim);
        for (size_t k = 0; k < cnt; ++k) idx[k] = ci(buf.data() + k * dim);
    }
    return (long)cnt;
}

template <typename T>
long refEvolveIter(const T* P, long n, T* S, long K, int dim, T r, int* counts, T* sums) {
    std::fill(sums, sums + K * dim, T(0));
    std::fill(counts, counts + K, 0);
    rd::detail::calculateClosestSphereCenter(P, S, sums, counts, (size_t)n, (size_t)K, (size_t)dim, r * r);
    int cont = 0;
    rd::detail::shiftTowardMassCenter(S, sums, counts, (size_t)K, (size_t)dim, cont);
    return cont;
}

template <typename T>
long refEvolve(const T* P, long n, T* S, long K, int dim, T r) {
    rd::evolve(P, S, (size_t)n, (size_t)K, (size_t)dim, r);
    return 0;
}

template <typename T>
long refEvolveOmpIter(const T* P, long n, T* S, long K, int dim, T r, int* counts, T* sums) {
    std::fill(sums, sums + K * dim, T(0));
    std::fill(counts, counts + K, 0);
    int threads = 1;
#if defined(RD_USE_OPENMP)
    threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
    {
        rd::detail::calculateClosestSphereCenter_omp(P, S, sums, counts, (size_t)n, (size_t)K,
                                                     (size_t)dim, r * r);
    }
    int cont = 0;
    rd::detail::shiftTowardMassCenter_omp(S, sums, counts, (size_t)K, (size_t)dim, cont, threads);
    return cont;
}

template <typename T>
long refDecimate(T* S, long K, int dim, T r2, long* kept) {
    std::vector<T> orig(S, S + K * dim);
    std::list<T*> lst;
    for (long k = 0; k < K; ++k) lst.push_back(S + k * dim);
    size_t cnt = (size_t)K;
    rd::decimate(S, lst, cnt, (size_t)dim, r2);
    if (kept) {
        CoordIndex<T> ci(orig.data(), K, dim);
        for (size_t k = 0; k < cnt; ++k) kept[k] = ci(S + k * dim);
    }
    return (long)cnt;
}

// decimateNoCopy over `nLists` consecutive groups of sizes listSizes[]; alive[] receives 1/0.
template <typename T>
long refDecimateNoCopy(T* S, long K, int dim, T r2, int nLists, const long* listSizes,
                       unsigned char* alive) {
    std::vector<std::list<T*>> lists((size_t)nLists);
    std::list<std::list<T*>*> g;
    long off = 0;
    for (int l = 0; l < nLists; ++l) {
        for (long k = 0; k < listSizes[l]; ++k) lists[l].push_back(S + (off + k) * dim);
        off += listSizes[l];
        g.push_back(&lists[l]);
    }
    size_t cnt = (size_t)K;
    rd::decimateNoCopy(g, cnt, (size_t)dim, r2);
    std::fill(alive, alive + K, 0);
    for (auto& l : lists)
        for (T* p : l) alive[(p - S) / dim] = 1;
    return (long)cnt;
}

template <typename T>
long refRD(const T* P, long n, int dim, T r1, T r2, T* S, int omp) {
    rd::RidgeDetection<T> rdc;
    if (!omp) rdc.noOMP();
    std::vector<T> buf((size_t)n * dim);
    rdc.ridgeDetection(P, (size_t)n, buf.data(), r1, r2, (size_t)dim);
    std::memcpy(S, buf.data(), sizeof(T) * rdc.ns_ * dim);
    return (long)rdc.ns_;
}

template <rd::tiled::TiledRDAlgorithm ALG, typename T>
long refTiledT(const T* P, long n, int dim, T r1, T r2, long cap, const long* tpd, int refinement,
               int threads, T* S, long Scap, float* timesMs) {
    rd::tiled::rdTiledData<T> dp;
    dp.dim = (size_t)dim;
    dp.np = (size_t)n;
    dp.maxTileCapacity = (size_t)cap;
    dp.nTilesPerDim.assign(tpd, tpd + dim);
    dp.r1 = r1;
    dp.r2 = r2;
    dp.P = P;
    auto t = rd::tiled::tiledRidgeDetection<ALG, T>(dp, threads, refinement != 0, false, false);
    if (timesMs) {
        timesMs[0] = std::get<0>(t);
        timesMs[1] = std::get<1>(t);
        timesMs[2] = std::get<2>(t);
        timesMs[3] = std::get<3>(t);
    }
    // The result is the first dataPack.ns rows of dataPack.S: the refinement's decimate
    // compacts S in place and lowers ns without shrinking the vector.
    long ns = (long)dp.ns;
    if (ns > Scap || (size_t)ns * dim > dp.S.size()) return -1;
    std::memcpy(S, dp.S.data(), sizeof(T) * (size_t)ns * dim);
    return ns;
}

template <typename T>
long refTiled(int alg, const T* P, long n, int dim, T r1, T r2, long cap, const long* tpd,
              int refinement, int threads, T* S, long Scap, float* timesMs) {
    using namespace rd::tiled;
    switch (alg) {
        case TILED_LOCAL_TREE_MIXED_RD:
            return refTiledT<TILED_LOCAL_TREE_MIXED_RD, T>(P, n, dim, r1, r2, cap, tpd, refinement, threads, S, Scap, timesMs);
        case TILED_GROUPED_TREE_MIXED_RD:
            return refTiledT<TILED_GROUPED_TREE_MIXED_RD, T>(P, n, dim, r1, r2, cap, tpd, refinement, threads, S, Scap, timesMs);
        case TILED_LOCAL_TREE_LOCAL_RD:
            return refTiledT<TILED_LOCAL_TREE_LOCAL_RD, T>(P, n, dim, r1, r2, cap, tpd, refinement, threads, S, Scap, timesMs);
        case TILED_LOCAL_LIST:
            return refTiledT<TILED_LOCAL_LIST, T>(P, n, dim, r1, r2, cap, tpd, refinement, threads, S, Scap, timesMs);
        case TILED_GROUPED_TREE_LOCAL_RD:
            return refTiledT<TILED_GROUPED_TREE_LOCAL_RD, T>(P, n, dim, r1, r2, cap, tpd, refinement, threads, S, Scap, timesMs);
        case EXT_TILE_TREE_MIXED_RD:
            return refTiledT<EXT_TILE_TREE_MIXED_RD, T>(P, n, dim, r1, r2, cap, tpd, refinement, threads, S, Scap, timesMs);
    }
    return -1;
}

// Flattened tree dump in the same format as oracle_tree_*.
template <typename Tree, typename T>
long dumpTree(Tree& tree, int kind, const T* P, long n, int dim, long* leafInfo, long infoCap,
              long* ownIdx, long ownCap, long* nbIdx, long nbCap, T* bounds) {
    CoordIndex<T> ci(P, n, dim);
    std::vector<typename Tree::SptrTile> leaves;
    tree.forEachLeaf([&](typename Tree::SptrTile t) { leaves.push_back(t); });
    long L = (long)leaves.size();
    if (!leafInfo) return L;
    if (L > infoCap) return -1;
    std::map<int, long> idToLeaf;
    for (long l = 0; l < L; ++l) idToLeaf[leaves[l]->id] = l;
    long o = 0, b = 0;
    for (long l = 0; l < L; ++l) {
        auto& t = *leaves[l];
        leafInfo[4 * l] = (long)t.pointsCnt;
        leafInfo[4 * l + 2] = (long)t.chosenPointsCapacity;
        leafInfo[4 * l + 3] = 0;
        if (o + (long)t.pointsCnt > ownCap) return -1;
        for (size_t i = 0; i < t.pointsCnt; ++i) ownIdx[o++] = ci(t.samples.data() + i * dim);
        for (int d = 0; d < dim; ++d) {
            bounds[(2 * l) * dim + 2 * d] = t.bounds.min(d);
            bounds[(2 * l) * dim + 2 * d + 1] = t.bounds.max(d);
        }
        leafInfo[4 * l + 1] = 0;
    }
    return L;
}

template <typename T>
long refTree(int kind, const T* P, long n, int dim, T r1, long cap, const long* tpd, long* leafInfo,
             long infoCap, long* ownIdx, long ownCap, long* nbIdx, long nbCap, T* bounds) {
    using namespace rd::tiled;
    std::vector<size_t> init(tpd, tpd + dim);
    CoordIndex<T> ci(P, n, dim);
    if (kind == 0) {
        TiledLocalTree<T> tree;
        tree.maxTileCapacity = (size_t)cap;
        tree.sphereRadius = r1;
        tree.buildTree(P, (size_t)n, (size_t)dim, init);
        return dumpTree(tree, kind, P, n, dim, leafInfo, infoCap, ownIdx, ownCap, nbIdx, nbCap, bounds);
    }
    if (kind == 1) {
        TiledGroupTree<FIND_NEIGHBOURS, T> tree;
        tree.maxTileCapacity = (size_t)cap;
        tree.sphereRadius = r1;
        tree.buildTree(P, (size_t)n, (size_t)dim, init);
        long L = dumpTree(tree, kind, P, n, dim, leafInfo, infoCap, ownIdx, ownCap, nbIdx, nbCap, bounds);
        if (L < 0 || !leafInfo) return L;
        std::vector<typename TiledGroupTree<FIND_NEIGHBOURS, T>::SptrTile> leaves;
        tree.forEachLeaf([&](typename TiledGroupTree<FIND_NEIGHBOURS, T>::SptrTile t) { leaves.push_back(t); });
        std::map<int, long> idToLeaf;
        for (long l = 0; l < L; ++l) idToLeaf[leaves[l]->id] = l;
        long b = 0;
        for (long l = 0; l < L; ++l) {
            leafInfo[4 * l + 1] = (long)leaves[l]->neighbours.size();
            if (b + (long)leaves[l]->neighbours.size() > nbCap) return -1;
            for (auto& nbt : leaves[l]->neighbours) {
                auto it = idToLeaf.find(nbt->id);
                nbIdx[b++] = it == idToLeaf.end() ? -1 : it->second;
            }
        }
        return L;
    }
    TiledGroupTree<EXTENDED_TILE, T> tree;
    tree.maxTileCapacity = (size_t)cap;
    tree.sphereRadius = r1;
    tree.buildTree(P, (size_t)n, (size_t)dim, init);
    long L = dumpTree(tree, kind, P, n, dim, leafInfo, infoCap, ownIdx, ownCap, nbIdx, nbCap, bounds);
    if (L < 0 || !leafInfo) return L;
    std::vector<typename TiledGroupTree<EXTENDED_TILE, T>::SptrTile> leaves;
    tree.forEachLeaf([&](typename TiledGroupTree<EXTENDED_TILE, T>::SptrTile t) { leaves.push_back(t); });
    long b = 0;
    for (long l = 0; l < L; ++l) {
        long cnt = (long)(leaves[l]->neighbours.size() / (size_t)dim);
        leafInfo[4 * l + 1] = cnt;
        if (b + cnt > nbCap) return -1;
        for (long i = 0; i < cnt; ++i) nbIdx[b++] = ci(leaves[l]->neighbours.data() + i * dim);
    }
    return L;
}

}  // namespace

// rd::orderSamples (rd/cpu/brute_force/order.hpp:42-142) with the condition of
// ridge_detection.inl:148-168; chains as point indices (pointer - S) / dim.
template <typename T>
long refOrder(const T* S, long K, int dim, T r2, long* order, long* offsets) {
    const T r2Sqr = r2 * r2;
    auto chains = rd::orderSamples<T>(S, (size_t)dim, (size_t)K, [&r2Sqr](T const* p1, T const* p2, size_t d) -> bool {
        return rd::squareEuclideanDistance(p1, p2, d) <= r2Sqr;
    });
    long nc = 0, w = 0;
    for (auto& c : chains) {
        offsets[nc++] = w;
        for (T const* p : c) order[w++] = (long)((p - S) / dim);
    }
    offsets[nc] = w;
    return nc;
}

// rd::RDAssessmentQuality (rd/utils/assessment_quality.hpp:91-220) with B as the ideal pattern.
template <typename T>
struct RefQuality : rd::RDAssessmentQuality<T> {
    RefQuality(const T* B, long nB, int dim) : rd::RDAssessmentQuality<T>((size_t)nB, (size_t)dim) {
        std::copy(B, B + nB * dim, this->idealPattern.begin());
    }
};

template <typename T>
long refHausdorff(const T* A, long nA, const T* B, long nB, int dim, T* out) {
    RefQuality<T> q(B, nB, dim);
    *out = q.hausdorffDistance(std::vector<T>(A, A + nA * dim));
    return 0;
}

template <typename T>
long refDistanceStats(const T* A, long nA, const T* B, long nB, int dim, T* out4) {
    RefQuality<T> q(B, nB, dim);
    q.setDistanceStats(std::vector<T>(A, A + nA * dim), out4[0], out4[1], out4[2], out4[3]);
    return 0;
}

extern "C" {

int ref_openmp_threads() {
#if defined(RD_USE_OPENMP)
    return omp_get_max_threads();
#else
    return 1;
#endif
}

#define REF_DEFINE(SUF, T)                                                                           \
    long ref_choose_##SUF(const T* P, long n, int dim, T r, T* S, long* idx) {                       \
        return refChoose<T>(P, n, dim, r, S, idx);                                                   \
    }                                                                                                \
    long ref_evolve_iter_##SUF(const T* P, long n, T* S, long K, int dim, T r, int* counts, T* sums) {\
        return refEvolveIter<T>(P, n, S, K, dim, r, counts, sums);                                   \
    }                                                                                                \
    long ref_evolve_##SUF(const T* P, long n, T* S, long K, int dim, T r) {                          \
        return refEvolve<T>(P, n, S, K, dim, r);                                                     \
    }                                                                                                \
    long ref_evolve_omp_iter_##SUF(const T* P, long n, T* S, long K, int dim, T r, int* counts,     \
                                   T* sums) {                                                        \
        return refEvolveOmpIter<T>(P, n, S, K, dim, r, counts, sums);                                \
    }                                                                                                \
    long ref_decimate_##SUF(T* S, long K, int dim, T r2, long* kept) {                               \
        return refDecimate<T>(S, K, dim, r2, kept);                                                  \
    }                                                                                                \
    long ref_decimate_nocopy_##SUF(T* S, long K, int dim, T r2, int nLists, const long* sizes,       \
                                   unsigned char* alive) {                                           \
        return refDecimateNoCopy<T>(S, K, dim, r2, nLists, sizes, alive);                            \
    }                                                                                                \
    long ref_hausdorff_##SUF(const T* A, long nA, const T* B, long nB, int dim, T* out) {            \
        return refHausdorff<T>(A, nA, B, nB, dim, out);                                              \
    }                                                                                                \
    long ref_distance_stats_##SUF(const T* A, long nA, const T* B, long nB, int dim, T* out4) {     \
        return refDistanceStats<T>(A, nA, B, nB, dim, out4);                                         \
    }                                                                                                \
    long ref_order_##SUF(const T* S, long K, int dim, T r2, long* order, long* offsets) {            \
        return refOrder<T>(S, K, dim, r2, order, offsets);                                           \
    }                                                                                                \
    long ref_rd_##SUF(const T* P, long n, int dim, T r1, T r2, T* S, int omp) {                      \
        return refRD<T>(P, n, dim, r1, r2, S, omp);                                                  \
    }                                                                                                \
    long ref_tiled_##SUF(int alg, const T* P, long n, int dim, T r1, T r2, long cap, const long* tpd,\
                         int refinement, int threads, T* S, long Scap, float* timesMs) {             \
        return refTiled<T>(alg, P, n, dim, r1, r2, cap, tpd, refinement, threads, S, Scap, timesMs); \
    }                                                                                                \
    long ref_tree_##SUF(int kind, const T* P, long n, int dim, T r1, long cap, const long* tpd,      \
                        long* leafInfo, long infoCap, long* ownIdx, long ownCap, long* nbIdx,        \
                        long nbCap, T* bounds) {                                                     \
        return refTree<T>(kind, P, n, dim, r1, cap, tpd, leafInfo, infoCap, ownIdx, ownCap, nbIdx,   \
                          nbCap, bounds);                                                            \
    }

REF_DEFINE(f32, float)
REF_DEFINE(f64, double)

}  // extern "C"
