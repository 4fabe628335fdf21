// Kernels shared by both precisions of the sift engine (ensemble mean,
// fills, Philox noise) and the host dispatchers of the per-precision sift
// kernels (eemd_sift_f32.hip, eemd_sift_f64.hip).
#include "eemd_sift_kernel.h"

namespace eemd {

// Level-wise ensemble mean and residual (emd_ref.py:209-217) of one sample:
// the task accumulators of its signal are added in task (= member) order.
// With `sums` the levels hold the sums (not divided by the member counts), for
// callers that combine several batches of one ensemble.
__device__ __forceinline__ void mean_sample(const SignalDesc* sig, int nsig, const int64_t* sig_start,
                                            const TaskDesc* task, const double* acc,
                                            const int32_t* acc_cnt, int max_imf, const double* x,
                                            double* avg, double* resid, int32_t* nmodes,
                                            int64_t gid, int sums) {
  int lo = 0, hi = nsig - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sig_start[mid] <= gid) lo = mid; else hi = mid - 1;
  }
  const SignalDesc sd = sig[lo];
  const int64_t i = gid - sig_start[lo];
  double rs = 0.0;
  int nm = 0;
  for (int m = 0; m < max_imf; ++m) {
    int cnt = 0;
    double s = 0.0;
    const int64_t mo = (int64_t)m * sd.n + i;
    // every task owns max_imf accumulator rows, so the value is loaded
    // unconditionally (independent loads, no wait on the count first) and a
    // level the task never reached is dropped by the select
#pragma unroll 8
    for (int t = sd.task0; t < sd.task0 + sd.ntask; ++t) {
      const int c = acc_cnt[(int64_t)t * max_imf + m];
      const double v = acc[task[t].acc_off + mo];
      s = (c > 0) ? s + v : s;
      cnt += c;
    }
    if (cnt == 0) break;
    const double a = sums ? s : s / (double)cnt;
    avg[sd.avg_off + (int64_t)m * sd.n + i] = a;
    rs = (m == 0) ? a : rs + a;
    nm = m + 1;
  }
  resid[sd.x_off + i] = x[sd.x_off + i] - (nm > 0 ? rs : 0.0);
  if (i == 0) nmodes[lo] = nm;
}

// One wave per MEAN_SPW consecutive samples, lane l taking samples l, l + 64,
// ... of them.  When they all belong to one signal (every wave of a long
// signal), the signal and the tasks' counts and accumulator offsets are the
// same for the whole wave and are read once into scalar registers, so each
// accumulator value costs one coalesced vector load; a wave that straddles
// signals runs the per-sample form.  The sums and their order are those of
// mean_sample.
constexpr int MEAN_SPL = 4;                   // samples per lane
constexpr int MEAN_SPW = MEAN_SPL * WAVE;     // samples per wave

__global__ void mean_kernel(const SignalDesc* sig, int nsig, const int64_t* sig_start,
                            const TaskDesc* task, const double* acc, const int32_t* acc_cnt,
                            int max_imf, const double* x, double* avg, double* resid,
                            int32_t* nmodes, int64_t total_n, int sums) {
  const int ln = (int)(threadIdx.x % WAVE);
  const int64_t base = ((int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * MEAN_SPW;
  if (base >= total_n) return;  // wave-uniform
  const int64_t wend = (base + MEAN_SPW < total_n) ? base + MEAN_SPW : total_n;
  int lo = 0, hi = nsig - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sig_start[mid] <= base) lo = mid; else hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  const SignalDesc sd = sig[lo];
  const int64_t s0 = sig_start[lo];
  if (wend > s0 + sd.n) {
#pragma unroll 1
    for (int k = 0; k < MEAN_SPL; ++k) {
      const int64_t gid = base + ln + (int64_t)k * WAVE;
      if (gid < total_n)
        mean_sample(sig, nsig, sig_start, task, acc, acc_cnt, max_imf, x, avg, resid, nmodes,
                    gid, sums);
    }
    return;
  }
  const int64_t i0 = base - s0 + ln;  // sample of k = 0 within the signal
  bool ok[MEAN_SPL];
#pragma unroll
  for (int k = 0; k < MEAN_SPL; ++k) ok[k] = base + ln + (int64_t)k * WAVE < wend;
  double rs[MEAN_SPL];
#pragma unroll
  for (int k = 0; k < MEAN_SPL; ++k) rs[k] = 0.0;
  int nm = 0;
  for (int m = 0; m < max_imf; ++m) {
    int cnt = 0;
    double s[MEAN_SPL];
#pragma unroll
    for (int k = 0; k < MEAN_SPL; ++k) s[k] = 0.0;
    const int64_t mo = (int64_t)m * sd.n + i0;
#pragma unroll 4
    for (int t = sd.task0; t < sd.task0 + sd.ntask; ++t) {
      const int c = acc_cnt[(int64_t)t * max_imf + m];
      const double* const a = acc + task[t].acc_off + mo;
#pragma unroll
      for (int k = 0; k < MEAN_SPL; ++k) {
        const double v = ok[k] ? a[k * WAVE] : 0.0;
        s[k] = (c > 0) ? s[k] + v : s[k];
      }
      cnt += c;
    }
    if (cnt == 0) break;  // wave-uniform
#pragma unroll
    for (int k = 0; k < MEAN_SPL; ++k) {
      const double av = sums ? s[k] : s[k] / (double)cnt;
      if (ok[k]) avg[sd.avg_off + mo + k * WAVE] = av;
      rs[k] = (m == 0) ? av : rs[k] + av;
    }
    nm = m + 1;
  }
#pragma unroll
  for (int k = 0; k < MEAN_SPL; ++k) {
    if (!ok[k]) continue;
    const int64_t q = sd.x_off + i0 + k * WAVE;
    resid[q] = x[q] - (nm > 0 ? rs[k] : 0.0);
  }
  if (base == s0 && ln == 0) nmodes[lo] = nm;
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_f64_kernel(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_bytes_kernel(uint8_t* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Materialise the in-kernel noise: z[e][i] for members [m0, m0+E).
__global__ void noise_kernel(uint64_t seed, uint32_t sig, uint32_t m0, int E, int64_t n,
                             double* out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // sample pair
  const int e = (int)blockIdx.y;
  if (e >= E || 2 * j >= n) return;
  double z0, z1;
  normal_pair(seed, sig, m0 + (uint32_t)e, (uint32_t)j, z0, z1);
  out[(int64_t)e * n + 2 * j] = z0;
  if (2 * j + 1 < n) out[(int64_t)e * n + 2 * j + 1] = z1;
}

// ------------------------------------------------------------------------
// host-side launchers (called from the C-ABI)
// ------------------------------------------------------------------------
hipError_t launch_sift_f32(const SiftArgs& A, int nt, int grid, hipStream_t st);
hipError_t launch_sift_f64(const SiftArgs& A, int nt, int grid, hipStream_t st);
hipError_t sift_occupancy_f32(int nt, int end_mode, int* blocks_per_cu);
hipError_t sift_occupancy_f64(int nt, int end_mode, int* blocks_per_cu);

hipError_t launch_sift(const SiftArgs& A, int fp64, int nt, int grid, hipStream_t st) {
  return fp64 ? launch_sift_f64(A, nt, grid, st) : launch_sift_f32(A, nt, grid, st);
}

hipError_t sift_occupancy(int fp64, int nt, int end_mode, int* blocks_per_cu) {
  return fp64 ? sift_occupancy_f64(nt, end_mode, blocks_per_cu)
              : sift_occupancy_f32(nt, end_mode, blocks_per_cu);
}

hipError_t launch_mean(const SignalDesc* sig, int nsig, const int64_t* sig_start,
                       const TaskDesc* task, const double* acc, const int32_t* acc_cnt,
                       int max_imf, const double* x, double* avg, double* resid,
                       int32_t* nmodes, int64_t total_n, int sums, hipStream_t st) {
  if (total_n <= 0) return hipSuccess;
  const int bs = 256;
  const int64_t per_block = (int64_t)(bs / WAVE) * MEAN_SPW;
  const int64_t g = (total_n + per_block - 1) / per_block;
  hipLaunchKernelGGL(mean_kernel, dim3((unsigned)g), dim3(bs), 0, st, sig, nsig, sig_start,
                     task, acc, acc_cnt, max_imf, x, avg, resid, nmodes, total_n, sums);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_f64(double* p, int64_t n, double v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_bytes(void* p, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_bytes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     (uint8_t*)p, n);
  return hipGetLastError();
}

hipError_t launch_noise(uint64_t seed, uint32_t sig, uint32_t m0, int E, int64_t n, double* out,
                        hipStream_t st) {
  if (E <= 0 || n <= 0) return hipSuccess;
  const int64_t pairs = (n + 1) / 2;
  hipLaunchKernelGGL(noise_kernel, dim3((unsigned)((pairs + 255) / 256), (unsigned)E), dim3(256),
                     0, st, seed, sig, m0, E, n, out);
  return hipGetLastError();
}

}  // namespace eemd
