// Sift kernels in FP32 with a workgroup per member (256, 512, 1024 threads),
// in a translation unit of their own so the build can give them their own
// scheduling flags; the one-wave kernel (64 threads) is in eemd_sift_f32w.hip.
#include "eemd_sift_kernel.h"

namespace eemd {

hipError_t launch_sift_f32w(const SiftArgs& A, int grid, hipStream_t st);
hipError_t sift_occupancy_f32w(int end_mode, int* blocks_per_cu);

hipError_t launch_sift_f32(const SiftArgs& A, int nt, int grid, hipStream_t st) {
  switch (nt) {
    case 64: return launch_sift_f32w(A, grid, st);
    case 256: return launch_sift_t<float, 256>(A, grid, st);
    case 512: return launch_sift_t<float, 512>(A, grid, st);
    case 1024: return launch_sift_t<float, 1024>(A, grid, st);
  }
  return hipErrorInvalidValue;
}

hipError_t sift_occupancy_f32(int nt, int end_mode, int* blocks_per_cu) {
  switch (nt) {
    case 64: return sift_occupancy_f32w(end_mode, blocks_per_cu);
    case 256: return sift_occupancy_t<float, 256>(end_mode, blocks_per_cu);
    case 512: return sift_occupancy_t<float, 512>(end_mode, blocks_per_cu);
    case 1024: return sift_occupancy_t<float, 1024>(end_mode, blocks_per_cu);
  }
  return hipErrorInvalidValue;
}

}  // namespace eemd
