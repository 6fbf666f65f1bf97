// rd_scan_f32.hip -- the evolve scan kernels for float (rd_scan_impl.hpp).
#include "rd_scan_impl.hpp"

namespace rdhip {

template <>
hipError_t launch_evolve_scan<float>(const ScanArgs<float>& a, int dim, int numItems, int ppt, int threads,
                                  int persistentBlocks, hipStream_t st) {
    if (a.xTable && a.mfma) return launch_evolve_scan_m(a, dim, numItems, ppt, threads, persistentBlocks, st);
    if (a.xTable) return launch_evolve_scan_x<float>(a, dim, numItems, ppt, threads, persistentBlocks, st);
    return launch_evolve_scan_impl<float, 0>(a, dim, numItems, ppt, threads, persistentBlocks, st);
}

}  // namespace rdhip
