// Sift kernels in FP64 (every workgroup size), in a translation unit of their
// own so the build can give them their own scheduling flags.
#include "eemd_sift_kernel.h"

namespace eemd {

hipError_t launch_sift_f64(const SiftArgs& A, int nt, int grid, hipStream_t st) {
  return launch_sift_prec<double>(A, nt, grid, st);
}

hipError_t sift_occupancy_f64(int nt, int end_mode, int* blocks_per_cu) {
  return sift_occupancy_prec<double>(nt, end_mode, blocks_per_cu);
}

}  // namespace eemd
