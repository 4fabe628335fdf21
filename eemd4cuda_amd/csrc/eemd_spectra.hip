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

// Spectra state a context owns: one grow-only device buffer carved into the
// arrays of a call, and the last batched FFT plan of each entry point (reused
// while rows and length repeat).  No allocation on repeated calls of one shape.
struct FftPlan {
  hipfftHandle h = 0;
  bool have = false;
  int64_t rows = 0, n = 0;
  void reset() {
    if (have) (void)hipfftDestroy(h);
    have = false;
  }
};

struct SpectraState {
  void* buf = nullptr;
  size_t bytes = 0;
  FftPlan hil, stft;
};

void spectra_free(void* p) {
  SpectraState* s = static_cast<SpectraState*>(p);
  s->hil.reset();
  s->stft.reset();
  if (s->buf) (void)hipFree(s->buf);
  delete s;
}

SpectraState* state(eemd_ctx* ctx) {
  void*& slot = ctx_spectra(ctx);
  if (!slot) {
    slot = new SpectraState();
    ctx_spectra_dtor(ctx, spectra_free);
  }
  return static_cast<SpectraState*>(slot);
}

hipError_t ensure(SpectraState* s, size_t want) {
  if (want <= s->bytes && s->buf) return hipSuccess;
  if (s->buf) {
    const hipError_t e = hipFree(s->buf);
    s->buf = nullptr;
    s->bytes = 0;
    if (e != hipSuccess) return e;
  }
  const hipError_t e = hipMalloc(&s->buf, want);
  if (e != hipSuccess) {
    s->buf = nullptr;
    return e;
  }
  s->bytes = want;
  return hipSuccess;
}

size_t al(size_t b) { return (b + 255) & ~size_t(255); }

int fail_hip(eemd_ctx* ctx, hipError_t e, const char* what) {
  (void)hipGetLastError();
  return ctx_fail(ctx, e == hipErrorOutOfMemory ? EEMD_ERR_OOM : EEMD_ERR_DEVICE,
                  (std::string(what) + ": " + hipGetErrorString(e)).c_str());
}

}  // namespace

extern "C" int eemd_hilbert_if(eemd_ctx* ctx, const double* x, int64_t rows, int64_t n, double sr,
                               double* inst_freq, double* inst_amp) {
  if (!ctx) return EEMD_ERR_BAD_ARG;
  if (!x || !inst_freq || !inst_amp || rows < 1 || n < 2 || !(sr > 0.0) || n > (int64_t(1) << 27) ||
      rows > INT_MAX)
    return ctx_fail(ctx, EEMD_ERR_BAD_ARG, "hilbert: bad arguments");
  hipError_t e = hipSetDevice(ctx_device(ctx));
  if (e != hipSuccess) return fail_hip(ctx, e, "set device");
  hipStream_t st = (hipStream_t)ctx_stream(ctx);
  SpectraState* S = state(ctx);
  const int64_t total = rows * n;
  const size_t bx = al(sizeof(double) * total), bz = al(sizeof(double2) * total);
  // x, z, phase, amplitude, frequency
  if ((e = ensure(S, bx + bz + 3 * bx)) != hipSuccess) return fail_hip(ctx, e, "hilbert scratch");
  char* base = static_cast<char*>(S->buf);
  double* dx = reinterpret_cast<double*>(base);
  double2* dz = reinterpret_cast<double2*>(base + bx);
  double* dph = reinterpret_cast<double*>(base + bx + bz);
  double* damp = reinterpret_cast<double*>(base + 2 * bx + bz);
  double* dif = reinterpret_cast<double*>(base + 3 * bx + bz);
  if (!S->hil.have || S->hil.rows != rows || S->hil.n != n) {
    S->hil.reset();
    int nn = (int)n;
    if (hipfftPlanMany(&S->hil.h, 1, &nn, nullptr, 1, nn, nullptr, 1, nn, HIPFFT_Z2Z, (int)rows) !=
        HIPFFT_SUCCESS)
      return ctx_fail(ctx, EEMD_ERR_DEVICE, "hilbert: hipfftPlanMany failed");
    S->hil.have = true;
    S->hil.rows = rows;
    S->hil.n = n;
  }
  if (hipfftSetStream(S->hil.h, st) != HIPFFT_SUCCESS)
    return ctx_fail(ctx, EEMD_ERR_DEVICE, "hilbert: hipfftSetStream failed");
  if ((e = hipMemcpyAsync(dx, x, sizeof(double) * total, hipMemcpyHostToDevice, st)) != hipSuccess)
    return fail_hip(ctx, e, "hilbert H2D");
  hipLaunchKernelGGL(real_to_complex_kernel, dim3(nblk(total, 256)), dim3(256), 0, st, dx, dz, total);
  if (hipfftExecZ2Z(S->hil.h, (hipfftDoubleComplex*)dz, (hipfftDoubleComplex*)dz, HIPFFT_FORWARD) !=
      HIPFFT_SUCCESS)
    return ctx_fail(ctx, EEMD_ERR_DEVICE, "hilbert: forward FFT failed");
  hipLaunchKernelGGL(analytic_mask_kernel, dim3(nblk(total, 256)), dim3(256), 0, st, dz, rows, n);
  if (hipfftExecZ2Z(S->hil.h, (hipfftDoubleComplex*)dz, (hipfftDoubleComplex*)dz, HIPFFT_BACKWARD) !=
      HIPFFT_SUCCESS)
    return ctx_fail(ctx, EEMD_ERR_DEVICE, "hilbert: inverse FFT failed");
  hipLaunchKernelGGL(angle_abs_kernel, dim3(nblk(total, 256)), dim3(256), 0, st, dz, dph, damp, total);
  hipLaunchKernelGGL(unwrap_kernel, dim3((unsigned)rows), dim3(1024), 0, st, dph, n);
  hipLaunchKernelGGL(gradient_kernel, dim3(nblk(total, 256)), dim3(256), 0, st, dph, dif, rows, n, sr);
  if ((e = hipGetLastError()) != hipSuccess) return fail_hip(ctx, e, "hilbert kernels");
  if ((e = hipMemcpyAsync(inst_freq, dif, sizeof(double) * total, hipMemcpyDeviceToHost, st)) !=
      hipSuccess ||
      (e = hipMemcpyAsync(inst_amp, damp, sizeof(double) * total, hipMemcpyDeviceToHost, st)) !=
      hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return fail_hip(ctx, e, "hilbert D2H");
  return EEMD_OK;
}

extern "C" int eemd_stft(eemd_ctx* ctx, const double* x, int64_t rows, int64_t n, int32_t nperseg,
                         int32_t noverlap, double* zxx) {
  if (!ctx) return EEMD_ERR_BAD_ARG;
  if (!x || !zxx || rows < 1 || n < 1 || nperseg < 2 || noverlap < 0 || noverlap >= nperseg)
    return ctx_fail(ctx, EEMD_ERR_BAD_ARG, "stft: bad arguments");
  const int L = nperseg, step = nperseg - noverlap, nf = L / 2 + 1;
  const int64_t len = n + 2 * (int64_t)(L / 2);
  if (len < L) return ctx_fail(ctx, EEMD_ERR_BAD_ARG, "stft: signal shorter than a frame");
  const int64_t nseg = (len - noverlap) / step;
  if (nseg < 1 || rows * nseg > (int64_t)INT32_MAX)
    return ctx_fail(ctx, EEMD_ERR_BAD_ARG, "stft: too many frames");
  hipError_t e = hipSetDevice(ctx_device(ctx));
  if (e != hipSuccess) return fail_hip(ctx, e, "set device");
  hipStream_t st = (hipStream_t)ctx_stream(ctx);
  SpectraState* S = state(ctx);
  double sumw = 0.0;
  for (int j = 0; j < L; ++j) sumw += 0.5 - 0.5 * std::cos(2.0 * 3.141592653589793 * j / (double)L);
  const int64_t nfr = rows * nseg;
  const size_t bx = al(sizeof(double) * rows * n), bfr = al(sizeof(double) * nfr * L),
               bsp = al(sizeof(double2) * nfr * nf);
  // x, frames, spectra, output
  if ((e = ensure(S, bx + bfr + 2 * bsp)) != hipSuccess) return fail_hip(ctx, e, "stft scratch");
  char* base = static_cast<char*>(S->buf);
  double* dx = reinterpret_cast<double*>(base);
  double* dfr = reinterpret_cast<double*>(base + bx);
  double2* dsp = reinterpret_cast<double2*>(base + bx + bfr);
  double2* dout = reinterpret_cast<double2*>(base + bx + bfr + bsp);
  if (!S->stft.have || S->stft.rows != nfr || S->stft.n != L) {
    S->stft.reset();
    int nn = L;
    if (hipfftPlanMany(&S->stft.h, 1, &nn, nullptr, 1, L, nullptr, 1, nf, HIPFFT_D2Z, (int)nfr) !=
        HIPFFT_SUCCESS)
      return ctx_fail(ctx, EEMD_ERR_DEVICE, "stft: hipfftPlanMany failed");
    S->stft.have = true;
    S->stft.rows = nfr;
    S->stft.n = L;
  }
  if (hipfftSetStream(S->stft.h, st) != HIPFFT_SUCCESS)
    return ctx_fail(ctx, EEMD_ERR_DEVICE, "stft: hipfftSetStream failed");
  if ((e = hipMemcpyAsync(dx, x, sizeof(double) * rows * n, hipMemcpyHostToDevice, st)) != hipSuccess)
    return fail_hip(ctx, e, "stft H2D");
  hipLaunchKernelGGL(stft_frames_kernel, dim3(nblk(nfr * L, 256)), dim3(256), 0, st, dx, dfr, rows,
                     n, nseg, L, step);
  if (hipfftExecD2Z(S->stft.h, (hipfftDoubleReal*)dfr, (hipfftDoubleComplex*)dsp) != HIPFFT_SUCCESS)
    return ctx_fail(ctx, EEMD_ERR_DEVICE, "stft: FFT failed");
  hipLaunchKernelGGL(stft_layout_kernel, dim3(nblk(nfr * nf, 256)), dim3(256), 0, st, dsp, dout,
                     rows, nseg, nf, 1.0 / sumw);
  if ((e = hipGetLastError()) != hipSuccess) return fail_hip(ctx, e, "stft kernels");
  if ((e = hipMemcpyAsync(zxx, dout, sizeof(double2) * nfr * nf, hipMemcpyDeviceToHost, st)) !=
      hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return fail_hip(ctx, e, "stft D2H");
  return EEMD_OK;
}
