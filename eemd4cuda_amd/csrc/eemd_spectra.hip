// Time-frequency views of IMFs on gfx950: Hilbert instantaneous frequency
// and amplitude, and the STFT (audio_eemd/spectra.py:18-48).
//
// Hilbert instantaneous frequency and amplitude
// (audio_eemd/spectra.py:32-48: z = imf + i H(imf) by scipy.signal.hilbert,
// inst_freq = np.gradient(np.unwrap(np.angle(z)), 1/sr) / 2 pi, inst_amp = |z|).
//
// The analytic signal uses hipFFT (batched Z2Z of length n, one row per
// IMF): spectrum times scipy's h (1 at DC and, for even n, Nyquist; 2 on
// positive frequencies; 0 on negative ones), inverse transform, 1/n.  Phase
// unwrapping is a per-row scan (one workgroup per row): numpy's correction
// ddmod - dd with ddmod = mod(dd + pi, 2 pi) - pi (pi kept for positive
// jumps of exactly pi) wherever |dd| >= pi, cumulated along the row.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <climits>
#include <cmath>
#include <string>

#include "../../include/eemd_mi355x.h"
#include "eemd_common.h"

namespace eemd {

__global__ void real_to_complex_kernel(const double* x, double2* z, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) z[i] = make_double2(x[i], 0.0);
}

// Z[r][k] *= h[k] / n
__global__ void analytic_mask_kernel(double2* Z, int64_t rows, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * n) return;
  const int64_t k = i % n;
  double h;
  if (k == 0) h = 1.0;
  else if ((n % 2 == 0) && k == n / 2) h = 1.0;
  else if (k < (n + 1) / 2) h = 2.0;
  else h = 0.0;
  const double s = h / (double)n;
  Z[i] = make_double2(Z[i].x * s, Z[i].y * s);
}

__global__ void angle_abs_kernel(const double2* z, double* phase, double* amp, int64_t total) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  phase[i] = atan2(z[i].y, z[i].x);
  amp[i] = hypot(z[i].x, z[i].y);
}

__device__ __forceinline__ double unwrap_corr(double prev, double cur) {
  const double pi = 3.141592653589793;
  const double dd = cur - prev;
  if (fabs(dd) < pi) return 0.0;
  double m = fmod(dd + pi, 2.0 * pi);
  if (m < 0.0) m += 2.0 * pi;
  double ddmod = m - pi;
  if (ddmod == -pi && dd > 0.0) ddmod = pi;
  return ddmod - dd;
}

// In place: p[r][i] += sum of the corrections of positions 1..i.
__global__ __launch_bounds__(1024) void unwrap_kernel(double* p, int64_t n) {
  __shared__ double part[1024];
  double* row = p + (int64_t)blockIdx.x * n;
  const int t = (int)threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t i0 = t * per, i1 = (i0 + per < n) ? i0 + per : n;
  double s = 0.0;
  for (int64_t i = (i0 < 1 ? 1 : i0); i < i1; ++i) s += unwrap_corr(row[i - 1], row[i]);
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan
    const double v = (t >= off) ? part[t - off] : 0.0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  double run = (t > 0) ? part[t - 1] : 0.0;
  // corrections depend on the original phases: read them before writing
  double prev = (i0 >= 1 && i0 < n) ? row[i0 - 1] : 0.0;
  __syncthreads();
  for (int64_t i = i0; i < i1; ++i) {
    const double cur = row[i];
    if (i >= 1) run += unwrap_corr(prev, cur);
    prev = cur;
    row[i] = cur + run;
  }
}

// np.gradient(f, 1/sr) / (2 pi): central differences inside, one-sided at the ends.
__global__ void gradient_kernel(const double* u, double* f, int64_t rows, int64_t n, double sr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * n) return;
  const int64_t k = i % n;
  const double dx = 1.0 / sr;
  double g;
  if (n == 1) g = 0.0;
  else if (k == 0) g = (u[i + 1] - u[i]) / dx;
  else if (k == n - 1) g = (u[i] - u[i - 1]) / dx;
  else g = (u[i + 1] - u[i - 1]) / (2.0 * dx);
  f[i] = g / (2.0 * 3.141592653589793);
}

// frames[r][k][j] = xpad[r][k*step + j] * w[j]; xpad = nperseg/2 zeros, x,
// nperseg/2 zeros (scipy.signal.stft boundary='zeros'); periodic Hann window.
__global__ void stft_frames_kernel(const double* x, double* frames, int64_t rows, int64_t n,
                                   int64_t nseg, int L, int step) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * nseg * L) return;
  const int j = (int)(i % L);
  const int64_t k = (i / L) % nseg;
  const int64_t r = i / ((int64_t)L * nseg);
  const int64_t src = k * step + j - L / 2;
  const double v = (src >= 0 && src < n) ? x[r * n + src] : 0.0;
  const double w = 0.5 - 0.5 * cos(2.0 * 3.141592653589793 * (double)j / (double)L);
  frames[i] = v * w;
}

// Zxx[r][f][k] = spec[r][k][f] * scale (complex, interleaved)
__global__ void stft_layout_kernel(const double2* spec, double2* out, int64_t rows, int64_t nseg,
                                   int nf, double scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * nseg * nf) return;
  const int64_t k = i % nseg;
  const int64_t f = (i / nseg) % nf;
  const int64_t r = i / (nseg * nf);
  const double2 v = spec[(r * nseg + k) * nf + f];
  out[i] = make_double2(v.x * scale, v.y * scale);
}

}  // namespace eemd

using namespace eemd;

namespace {

unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

extern "C" int eemd_hilbert_if(eemd_ctx* ctx, const double* x, int64_t rows, int64_t n, double sr,
                               double* inst_freq, double* inst_amp) {
  if (!ctx) return EEMD_ERR_BAD_ARG;
  if (!x || !inst_freq || !inst_amp || rows < 1 || n < 2 || !(sr > 0.0) || n > (int64_t(1) << 27))
    return EEMD_ERR_BAD_ARG;
  if (hipSetDevice(ctx_device(ctx)) != hipSuccess) {
    (void)hipGetLastError();
    return EEMD_ERR_DEVICE;
  }
  const int64_t total = rows * n;
  double *dx = nullptr, *dph = nullptr, *damp = nullptr, *dif = nullptr;
  double2* dz = nullptr;
  hipfftHandle plan = 0;
  bool have_plan = false;
  int rc = EEMD_OK;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && rc == EEMD_OK) rc = (e == hipErrorOutOfMemory) ? EEMD_ERR_OOM : EEMD_ERR_DEVICE;
    return e == hipSuccess;
  };
  if (chk(hipMalloc(&dx, sizeof(double) * total)) && chk(hipMalloc(&dz, sizeof(double2) * total)) &&
      chk(hipMalloc(&dph, sizeof(double) * total)) && chk(hipMalloc(&damp, sizeof(double) * total)) &&
      chk(hipMalloc(&dif, sizeof(double) * total)) &&
      chk(hipMemcpy(dx, x, sizeof(double) * total, hipMemcpyHostToDevice))) {
    int nn = (int)n;
    if (hipfftPlanMany(&plan, 1, &nn, nullptr, 1, nn, nullptr, 1, nn, HIPFFT_Z2Z, (int)rows) !=
        HIPFFT_SUCCESS) {
      rc = EEMD_ERR_DEVICE;
    } else {
      have_plan = true;
      hipLaunchKernelGGL(real_to_complex_kernel, dim3(nblk(total, 256)), dim3(256), 0, 0, dx, dz, total);
      if (hipfftExecZ2Z(plan, (hipfftDoubleComplex*)dz, (hipfftDoubleComplex*)dz, HIPFFT_FORWARD) !=
          HIPFFT_SUCCESS)
        rc = EEMD_ERR_DEVICE;
      if (rc == EEMD_OK) {
        hipLaunchKernelGGL(analytic_mask_kernel, dim3(nblk(total, 256)), dim3(256), 0, 0, dz, rows, n);
        if (hipfftExecZ2Z(plan, (hipfftDoubleComplex*)dz, (hipfftDoubleComplex*)dz,
                          HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
          rc = EEMD_ERR_DEVICE;
      }
      if (rc == EEMD_OK) {
        hipLaunchKernelGGL(angle_abs_kernel, dim3(nblk(total, 256)), dim3(256), 0, 0, dz, dph, damp, total);
        hipLaunchKernelGGL(unwrap_kernel, dim3((unsigned)rows), dim3(1024), 0, 0, dph, n);
        hipLaunchKernelGGL(gradient_kernel, dim3(nblk(total, 256)), dim3(256), 0, 0, dph, dif, rows, n, sr);
        if (chk(hipGetLastError()) && chk(hipDeviceSynchronize())) {
          chk(hipMemcpy(inst_freq, dif, sizeof(double) * total, hipMemcpyDeviceToHost));
          chk(hipMemcpy(inst_amp, damp, sizeof(double) * total, hipMemcpyDeviceToHost));
        }
      }
    }
  }
  if (have_plan) hipfftDestroy(plan);
  for (void* p : {(void*)dx, (void*)dz, (void*)dph, (void*)damp, (void*)dif})
    if (p) (void)hipFree(p);
  (void)hipGetLastError();
  return rc;
}

extern "C" int eemd_stft(eemd_ctx* ctx, const double* x, int64_t rows, int64_t n, int32_t nperseg,
                         int32_t noverlap, double* zxx) {
  if (!ctx) return EEMD_ERR_BAD_ARG;
  if (!x || !zxx || rows < 1 || n < 1 || nperseg < 2 || noverlap < 0 || noverlap >= nperseg)
    return EEMD_ERR_BAD_ARG;
  const int L = nperseg, step = nperseg - noverlap, nf = L / 2 + 1;
  const int64_t len = n + 2 * (int64_t)(L / 2);
  if (len < L) return EEMD_ERR_BAD_ARG;
  const int64_t nseg = (len - noverlap) / step;
  if (nseg < 1 || rows * nseg > (int64_t)INT32_MAX) return EEMD_ERR_BAD_ARG;
  if (hipSetDevice(ctx_device(ctx)) != hipSuccess) {
    (void)hipGetLastError();
    return EEMD_ERR_DEVICE;
  }
  double sumw = 0.0;
  for (int j = 0; j < L; ++j) sumw += 0.5 - 0.5 * std::cos(2.0 * 3.141592653589793 * j / (double)L);
  double *dx = nullptr, *dfr = nullptr;
  double2 *dsp = nullptr, *dout = nullptr;
  hipfftHandle plan = 0;
  bool have_plan = false;
  int rc = EEMD_OK;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && rc == EEMD_OK) rc = (e == hipErrorOutOfMemory) ? EEMD_ERR_OOM : EEMD_ERR_DEVICE;
    return e == hipSuccess;
  };
  const int64_t nfr = rows * nseg;
  if (chk(hipMalloc(&dx, sizeof(double) * rows * n)) &&
      chk(hipMalloc(&dfr, sizeof(double) * nfr * L)) &&
      chk(hipMalloc(&dsp, sizeof(double2) * nfr * nf)) &&
      chk(hipMalloc(&dout, sizeof(double2) * nfr * nf)) &&
      chk(hipMemcpy(dx, x, sizeof(double) * rows * n, hipMemcpyHostToDevice))) {
    int nn = L;
    if (hipfftPlanMany(&plan, 1, &nn, nullptr, 1, L, nullptr, 1, nf, HIPFFT_D2Z, (int)nfr) !=
        HIPFFT_SUCCESS) {
      rc = EEMD_ERR_DEVICE;
    } else {
      have_plan = true;
      hipLaunchKernelGGL(stft_frames_kernel, dim3(nblk(nfr * L, 256)), dim3(256), 0, 0, dx, dfr,
                         rows, n, nseg, L, step);
      if (hipfftExecD2Z(plan, (hipfftDoubleReal*)dfr, (hipfftDoubleComplex*)dsp) != HIPFFT_SUCCESS)
        rc = EEMD_ERR_DEVICE;
      if (rc == EEMD_OK) {
        hipLaunchKernelGGL(stft_layout_kernel, dim3(nblk(nfr * nf, 256)), dim3(256), 0, 0, dsp, dout,
                           rows, nseg, nf, 1.0 / sumw);
        if (chk(hipGetLastError()) && chk(hipDeviceSynchronize()))
          chk(hipMemcpy(zxx, dout, sizeof(double2) * nfr * nf, hipMemcpyDeviceToHost));
      }
    }
  }
  if (have_plan) hipfftDestroy(plan);
  for (void* p : {(void*)dx, (void*)dfr, (void*)dsp, (void*)dout})
    if (p) (void)hipFree(p);
  (void)hipGetLastError();
  return rc;
}
