// rd_tiled.hpp -- tiled ridge detection on the device (tree build + segment drivers).
#pragma once

#include "rd_engine.hpp"

#include <vector>

namespace rdhip {

template <typename T>
struct TiledParams {
    int algorithm = 5;
    T r1 = 0, r2 = 0;
    int64_t maxTileCapacity = 0;
    std::vector<int64_t> tilesPerDim;
    bool refinement = false;
    int64_t maxEvolveIters = 0;
};

template <typename T>
void tiled_ridge_detection(Engine<T>& e, const TiledParams<T>& p);

}  // namespace rdhip
