// rd_scan_f64.hip -- the evolve scan kernels for double (rd_scan_impl.hpp).
#include "rd_scan_impl.hpp"

namespace rdhip {

template <>
hipError_t launch_evolve_scan<double>(const ScanArgs<double>& a, int dim, int numItems, int ppt, int threads,
                                  int persistentBlocks, hipStream_t st) {
    if (a.mfma) return hipErrorInvalidValue;   // the matrix-core screen is FP32 only
    if (a.xTable) return launch_evolve_scan_x<double>(a, dim, numItems, ppt, threads, persistentBlocks, st);
    return launch_evolve_scan_impl<double, 0>(a, dim, numItems, ppt, threads, persistentBlocks, st);
}

}  // namespace rdhip
