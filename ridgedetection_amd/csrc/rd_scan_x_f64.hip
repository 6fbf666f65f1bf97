// rd_scan_x_f64.hip -- the evolve scan kernels for double with the expanded-form screen
// (rd_scan_impl.hpp, rd_nearest.hpp).
#include "rd_scan_impl.hpp"

namespace rdhip {

template <>
hipError_t launch_evolve_scan_x<double>(const ScanArgs<double>& a, int dim, int numItems, int ppt, int threads,
                                        int persistentBlocks, hipStream_t st) {
    return launch_evolve_scan_impl<double, 1>(a, dim, numItems, ppt, threads, persistentBlocks, st);
}

}  // namespace rdhip
