// rd_tiled.hip -- tiled ridge detection: tree on the device, then every leaf is a segment.
//
// Reference drivers (rd/cpu/tiled/rd_tiled.hpp):
//   TILED_LOCAL_TREE_MIXED_RD   :756-966   local tree;   choose per leaf, rounds of
//                                          { evolve per leaf over its own samples,
//                                            global decimateNoCopy }
//   TILED_GROUPED_TREE_MIXED_RD :968-1214  grouped tree; evolve over own + neighbour tiles
//   EXT_TILE_TREE_MIXED_RD      :1217-1438 extended tiles; evolve over own + nearby samples
//   TILED_LOCAL_TREE_LOCAL_RD   :1441-1578 a whole ridge detection per leaf (own samples)
//   TILED_GROUPED_TREE_LOCAL_RD :1592-1732 the same over own + neighbour tiles
//   TILED_LOCAL_LIST            :1580-1590 a stub with an empty result
// plus local RD over extended tiles.  Results are collected in leaf preorder; the optional
// refinement decimates the collected set and evolves it over every input sample
// (:919-948).
#include "rd_tiled.hpp"
#include "rd_tree.hpp"

#include <numeric>

namespace rdhip {

template <typename T>
void tiled_ridge_detection(Engine<T>& e, const TiledParams<T>& p) {
    auto t0 = std::chrono::steady_clock::now();
    RunOptions opt = e.options;
    opt.maxEvolveIters = p.maxEvolveIters;
    if (p.algorithm == 3) {   // TILED_LOCAL_LIST: the reference returns nothing
        e.setSegments({}, {}, {}, {}, false);
        return;
    }
    int kind = TREE_LOCAL;
    bool mixed = true;
    switch (p.algorithm) {
        case 0: kind = TREE_LOCAL; mixed = true; break;
        case 1: kind = TREE_GROUPED; mixed = true; break;
        case 2: kind = TREE_LOCAL; mixed = false; break;
        case 4: kind = TREE_GROUPED; mixed = false; break;
        case 5: kind = TREE_EXTENDED; mixed = true; break;
        case 6: kind = TREE_EXTENDED; mixed = false; break;
        default: throw std::invalid_argument("unknown tiled algorithm");
    }
    TileTreeBuilder<T> tb(e.dim(), e.stream());
    tb.build(e.devicePoints(), e.pointStride(), e.numPoints(), kind, p.r1, p.maxTileCapacity, p.tilesPerDim);
    const auto& leaves = tb.leaves();
    const int L = (int)leaves.size();
    std::vector<SegDesc> segs(L);
    std::vector<int64_t> caps(L);
    int64_t pos = 0;
    for (int l = 0; l < L; ++l) {
        SegDesc s{};
        s.pOff = pos;
        s.nOwn = (int32_t)leaves[l].ownCnt;
        s.nP = (int32_t)(leaves[l].ownCnt + leaves[l].nbCnt);
        s.group = mixed ? 0 : l;
        segs[l] = s;
        caps[l] = leaves[l].ownCnt;   // choose picks at most every own sample
        pos += s.nP;
    }
    int* idx = e.reserveIndex((size_t)std::max<int64_t>(tb.totalPositions(), 1));
    tb.writeIndex(idx);
    e.timers.buildMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::vector<T> r1(L, p.r1), r2(L, p.r2);
    e.setSegments(segs, caps, r1, r2, true);
    if (mixed)
        e.ridgeDetectionMixed(opt);
    else
        e.ridgeDetectionLocal(opt);
    if (p.refinement) {
        auto t1 = std::chrono::steady_clock::now();
        const int64_t K = e.totalCount();
        std::vector<T> S((size_t)std::max<int64_t>(K, 1) * e.dim());
        if (K > 0) e.getChosen(S.data(), LAYOUT_ROW_MAJOR, false);
        SegDesc s{};
        s.pOff = 0;
        s.nP = (int32_t)e.numPoints();
        s.nOwn = (int32_t)e.numPoints();
        s.group = 0;
        e.setSegments({s}, {std::max<int64_t>(K, 1)}, {p.r1}, {p.r2}, false);
        e.setChosen(0, S.data(), K, LAYOUT_ROW_MAJOR);
        e.decimate(false);
        e.evolve(p.maxEvolveIters);
        e.timers.refineMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
    }
    e.timers.wholeMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

template void tiled_ridge_detection<float>(Engine<float>&, const TiledParams<float>&);
template void tiled_ridge_detection<double>(Engine<double>&, const TiledParams<double>&);

}  // namespace rdhip
