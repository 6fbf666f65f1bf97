// rd_scan_x_f32.hip -- the evolve scan kernels for float with the expanded-form screen
// (rd_scan_impl.hpp, rd_nearest.hpp).
#include "rd_scan_impl.hpp"

namespace rdhip {

template <>
hipError_t launch_evolve_scan_x<float>(const ScanArgs<float>& a, int dim, int numItems, int ppt, int threads,
                                       int persistentBlocks, hipStream_t st) {
    return launch_evolve_scan_impl<float, 1>(a, dim, numItems, ppt, threads, persistentBlocks, st);
}

}  // namespace rdhip
