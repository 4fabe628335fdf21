// The FP32 sift kernel with one wave per member (64-thread workgroups, the
// shape of large batches), in a translation unit of its own: it keeps the
// default instruction scheduler, which runs its learned-ends path faster
// than the ILP-maximising one used for the other FP32 kernels.
#include "eemd_sift_kernel.h"

namespace eemd {

hipError_t launch_sift_f32w(const SiftArgs& A, int grid, hipStream_t st) {
  return launch_sift_t<float, 64>(A, grid, st);
}

hipError_t sift_occupancy_f32w(int end_mode, int* blocks_per_cu) {
  return sift_occupancy_t<float, 64>(end_mode, blocks_per_cu);
}

}  // namespace eemd
