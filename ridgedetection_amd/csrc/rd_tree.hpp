// rd_tree.hpp -- tile trees built on the device (local, grouped, extended tiles).
//
// Reference semantics (rd/cpu/tiled/local_tile_tree.inl:48-437,
// rd/cpu/tiled/grouped_tile_tree.inl:86-624, rd/utils/bounding_box.hpp, histogram.hpp):
//   * the root is split into a grid of tilesPerDim cells over the global tight bounding box;
//   * LOCAL tiles bin points with the histogram map (floor of the offset over the bin width,
//     last bin for the max) and then take the tight bounds of their own points;
//     GROUPED / EXTENDED tiles are grid cells (the last cell of a dimension ends at the
//     parent's max) and take the points inside the half-open cell;
//   * EXTENDED tiles also keep the samples of (parent own, parent neighbours) that are
//     near the cell (inside the r1-expanded cell but not inside it), in that order;
//   * tiles with 0 or 1 own points are dropped; a tile with at least maxTileCapacity own
//     points is split in two along dimension (level % dim), the root's children being
//     level 1; leaves are listed in preorder;
//   * GROUPED leaves' neighbours are the other leaves whose (closed) cell bounds overlap.
// Point lists keep input order at every level because every partition is a stable sort.
#pragma once

#include "rd_engine.hpp"

#include <memory>
#include <vector>

namespace rdhip {

enum TreeKind { TREE_LOCAL = 0, TREE_GROUPED = 1, TREE_EXTENDED = 2 };

template <typename T>
struct LeafInfo {
    int64_t ownCnt = 0;
    int64_t nbCnt = 0;             // EXTENDED: neighbour samples; GROUPED: samples of neighbour leaves
    std::vector<T> mn, mx;         // bounds (tight for LOCAL, cell for GROUPED / EXTENDED)
    size_t sphereCap = 0;          // the reference's chosen capacity (countSpheresInside(r1))
    std::vector<int> nbLeaves;     // GROUPED: neighbouring leaves (preorder positions)
};

template <typename T>
class TileTreeBuilder {
public:
    TileTreeBuilder(int dim, hipStream_t st) : dim_(dim), st_(st) {}
    ~TileTreeBuilder();

    void build(const T* P, int64_t pStride, int64_t n, int kind, T r1, int64_t maxTileCapacity,
               const std::vector<int64_t>& tilesPerDim);

    const std::vector<LeafInfo<T>>& leaves() const { return leaves_; }
    // Positions of the segment layout: per leaf, own samples then neighbour samples.
    int64_t totalPositions() const { return totalPositions_; }
    // Writes the point index of every position (device buffer of totalPositions()).
    void writeIndex(int* dIndex);
    int levels() const { return levels_; }

private:
    struct HTile {
        std::vector<T> mn, mx, dist;
        int level = 0;
        int ownBuf = -1;
        int64_t ownOff = 0, ownCnt = 0;
        int nbBuf = -1;
        int64_t nbOff = 0, nbCnt = 0;
        std::vector<int> children;
        bool leaf = false;
    };
    int dim_;
    hipStream_t st_;
    int kind_ = 0;
    T r1_ = 0;
    int64_t cap_ = 0;
    std::vector<HTile> tiles_;
    std::vector<int> rootChildren_;
    std::vector<std::unique_ptr<DevBuf<int>>> bufs_;
    std::vector<int> leafTiles_;
    std::vector<LeafInfo<T>> leaves_;
    int64_t totalPositions_ = 0;
    int levels_ = 0;

    void preorder(int t);
};

}  // namespace rdhip
