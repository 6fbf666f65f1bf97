// rd_tiled.hip -- tiled ridge detection (placeholder until the device tree build lands).
#include "rd_tiled.hpp"

namespace rdhip {

template <typename T>
void tiled_ridge_detection(Engine<T>&, const TiledParams<T>&) {
    throw std::runtime_error("tiled ridge detection is not built yet");
}

template void tiled_ridge_detection<float>(Engine<float>&, const TiledParams<float>&);
template void tiled_ridge_detection<double>(Engine<double>&, const TiledParams<double>&);

}  // namespace rdhip
