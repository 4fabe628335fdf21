// Sift kernels in FP64 (every workgroup size), in a translation unit of their
// own so the build can give them their own scheduling flags.
#include "eemd_sift_kernel.h"

namespace eemd {

hipError_t launch_sift_f64(const SiftArgs& A, int nt, int grid, hipStream_t st) {
  switch (nt) {
    case 64: return launch_sift_t<double, 64>(A, grid, st);
    case 256: return launch_sift_t<double, 256>(A, grid, st);
    case 512: return launch_sift_t<double, 512>(A, grid, st);
    case 1024: return launch_sift_t<double, 1024>(A, grid, st);
  }
  return hipErrorInvalidValue;
}

hipError_t sift_occupancy_f64(int nt, int end_mode, int* blocks_per_cu) {
  switch (nt) {
    case 64: return sift_occupancy_t<double, 64>(end_mode, blocks_per_cu);
    case 256: return sift_occupancy_t<double, 256>(end_mode, blocks_per_cu);
    case 512: return sift_occupancy_t<double, 512>(end_mode, blocks_per_cu);
    case 1024: return sift_occupancy_t<double, 1024>(end_mode, blocks_per_cu);
  }
  return hipErrorInvalidValue;
}

}  // namespace eemd
