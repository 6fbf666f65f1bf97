// rd_scan_m_f32.hip -- the evolve scan kernels for float with the expanded-form screen on the
// matrix cores (rd_scan_impl.hpp, nearest_labels_mfma in rd_nearest.hpp).
#include "rd_scan_impl.hpp"

namespace rdhip {

hipError_t launch_evolve_scan_m(const ScanArgs<float>& a, int dim, int numItems, int ppt, int threads,
                                int persistentBlocks, hipStream_t st) {
    return launch_evolve_scan_impl<float, 2>(a, dim, numItems, ppt, threads, persistentBlocks, st);
}

}  // namespace rdhip
