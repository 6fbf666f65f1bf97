// rd_scan_f32.hip -- the evolve scan kernels for float (rd_scan_impl.hpp).
#include "rd_scan_impl.hpp"

namespace rdhip {

template <>
hipError_t launch_evolve_scan<float>(const ScanArgs<float>& a, int dim, int numItems, int ppt, int threads,
                                  int persistentBlocks, hipStream_t st) {
    return launch_evolve_scan_impl<float>(a, dim, numItems, ppt, threads, persistentBlocks, st);
}

}  // namespace rdhip
