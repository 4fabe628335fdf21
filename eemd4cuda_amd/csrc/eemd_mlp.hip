// Training and batch inference of the learned boundary-extension MLP
// (64 -> 128 -> 128 -> 128 -> 16, GELU) on gfx950.
//
// Training keeps FP32 master parameters, Adam moments and a BF16 copy of the
// parameters.  Every product is a BF16 x BF16 -> FP32 MFMA GEMM
// (v_mfma_f32_16x16x16_bf16, one 16 x 16 output tile per wave): forward
// Z = X W + b, backward dX = dZ W^T and dW = X^T dZ, with the FP32 operands
// rounded to BF16 as they are loaded.  Loss: mean squared error over the
// 16 outputs.  Inference (eemd_mlp_infer and the in-kernel inference of the
// sift engine) uses the exact FP32 arithmetic of eemd_mlp.h.
//
// No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/eemd_mi355x.h"
#include "eemd_mlp.h"

namespace eemd {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t as_bf16(float v) { return f32_to_bf16(v); }
__device__ __forceinline__ uint16_t as_bf16(uint16_t v) { return v; }

// C[M][N] (ldc) = op(A) op(B) (+ bias[n] in BF16).  op(A)(m,k) = A[m*lda+k]
// or, transposed, A[k*lda+m]; op(B)(k,n) = B[k*ldb+n] or B[n*ldb+k].
// M, N, K multiples of 16.  MFMA 16x16x16 operand layout: lane l holds
// A(l%16, 4(l/16)+i) and B(4(l/16)+i, l%16), i = 0..3, and receives
// D(4(l/16)+i, l%16).
template <typename TA, typename TB>
__global__ __launch_bounds__(64) void gemm_bf16_kernel(const TA* A, int lda, int ta, const TB* B,
                                                       int ldb, int tb, float* C, int ldc,
                                                       const uint16_t* bias, int K) {
  const int n0 = blockIdx.x * 16, m0 = blockIdx.y * 16;
  const int l = threadIdx.x;
  const int r = l & 15, q = l >> 4;
  f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4_t a, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = k0 + 4 * q + i;
      const TA av = ta ? A[(int64_t)k * lda + m0 + r] : A[(int64_t)(m0 + r) * lda + k];
      const TB bv = tb ? B[(int64_t)(n0 + r) * ldb + k] : B[(int64_t)k * ldb + n0 + r];
      a[i] = (short)as_bf16(av);
      b[i] = (short)as_bf16(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + 4 * q + i, n = n0 + r;
    float v = acc[i];
    if (bias) v += bf16_to_f32(bias[n]);
    C[(int64_t)m * ldc + n] = v;
  }
}

__global__ void gelu_fwd_kernel(const float* z, float* a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = mlp_gelu(z[i]);
}

// dZ = dA * GELU'(Z) for the tanh form of mlp_gelu.
__global__ void gelu_bwd_kernel(const float* da, const float* z, float* dz, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = z[i];
  const float c = 0.7978845608028654f;
  const float u = c * (x + 0.044715f * x * x * x);
  const float t = tanhf(u);
  const float d = 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * x * x);
  dz[i] = da[i] * d;
}

// dZ = 2 (Yhat - Y) / n and the loss sum(Yhat - Y)^2 / n (one workgroup,
// fixed-order tree, so the loss is reproducible).
__global__ __launch_bounds__(256) void mse_kernel(const float* yh, const float* y, float* dz,
                                                  float* loss, int64_t n) {
  __shared__ float part[256];
  float s = 0.f;
  const float inv = 1.0f / (float)n;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const float d = yh[i] - y[i];
    dz[i] = 2.0f * d * inv;
    s += d * d;
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = part[0] * inv;
}

// db[n] = sum_b dZ[b][n]
__global__ void bias_grad_kernel(const float* dz, float* db, int batch, int nout) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nout) return;
  float s = 0.f;
  for (int b = 0; b < batch; ++b) s += dz[(int64_t)b * nout + j];
  db[j] = s;
}

__global__ void adam_kernel(float* th, const float* g, float* m, float* v, uint16_t* w16, int n,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float t = th[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  th[i] = t;
  w16[i] = f32_to_bf16(t);
}

__global__ void to_bf16_kernel(const float* th, uint16_t* w16, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w16[i] = f32_to_bf16(th[i]);
}

// Exact inference, one workgroup of 128 threads per sample.
__global__ __launch_bounds__(MLP_HID) void mlp_infer_kernel(const uint16_t* w, const float* X,
                                                            float* Y, int batch) {
  __shared__ float xin[MLP_IN], h0[MLP_HID], h1[MLP_HID];
  const int b = blockIdx.x;
  const int j = threadIdx.x;
  if (b >= batch) return;
  if (j < MLP_IN) xin[j] = X[(int64_t)b * MLP_IN + j];
  __syncthreads();
  h0[j] = mlp_gelu(mlp_neuron(w + MLP_W1 + j, MLP_HID, bf16_to_f32(w[MLP_B1 + j]), xin, MLP_IN));
  __syncthreads();
  h1[j] = mlp_gelu(mlp_neuron(w + MLP_W2 + j, MLP_HID, bf16_to_f32(w[MLP_B2 + j]), h0, MLP_HID));
  __syncthreads();
  h0[j] = mlp_gelu(mlp_neuron(w + MLP_W3 + j, MLP_HID, bf16_to_f32(w[MLP_B3 + j]), h1, MLP_HID));
  __syncthreads();
  if (j < MLP_OUT)
    Y[(int64_t)b * MLP_OUT + j] =
        mlp_neuron(w + MLP_W4 + j, MLP_OUT, bf16_to_f32(w[MLP_B4 + j]), h0, MLP_HID);
}

static_assert(MLP_NPARAM == EEMD_MLP_NPARAM, "parameter count of the C ABI");

}  // namespace eemd

using namespace eemd;

namespace {

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t want) {
    if (p && want <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    else p = nullptr;
    return e;
  }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

const int kIn[4] = {MLP_IN, MLP_HID, MLP_HID, MLP_HID};
const int kOut[4] = {MLP_HID, MLP_HID, MLP_HID, MLP_OUT};
const int kW[4] = {MLP_W1, MLP_W2, MLP_W3, MLP_W4};
const int kB[4] = {MLP_B1, MLP_B2, MLP_B3, MLP_B4};

}  // namespace

struct eemd_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  Buf th, g, m, v, w16, loss;
  Buf z[4], a[3], dz[4], da[3];  // activations of the last forward
  int cap = 0;                   // batch the activation buffers hold
  int t = 0;                     // Adam steps taken
  std::string err;
};

namespace {

int mlp_fail(eemd_mlp* h, hipError_t e, const char* what) {
  if (h) h->err = std::string(what) + ": " + hipGetErrorString(e);
  (void)hipGetLastError();
  return e == hipErrorOutOfMemory ? EEMD_ERR_OOM : EEMD_ERR_DEVICE;
}

#define MLPCHK(h, expr, what)                               \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return mlp_fail(h, _e, what);     \
  } while (0)

template <typename TA, typename TB>
hipError_t gemm(const TA* A, int lda, int ta, const TB* B, int ldb, int tb, float* C, int ldc,
                const uint16_t* bias, int M, int N, int K, hipStream_t st) {
  hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB>), dim3(N / 16, M / 16), dim3(64), 0, st, A, lda, ta,
                     B, ldb, tb, C, ldc, bias, K);
  return hipGetLastError();
}

unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t ensure_batch(eemd_mlp* h, int batch) {
  if (batch <= h->cap) return hipSuccess;
  hipError_t e;
  for (int l = 0; l < 4; ++l) {
    if ((e = h->z[l].ensure(sizeof(float) * batch * kOut[l])) != hipSuccess) return e;
    if ((e = h->dz[l].ensure(sizeof(float) * batch * kOut[l])) != hipSuccess) return e;
    if (l < 3) {
      if ((e = h->a[l].ensure(sizeof(float) * batch * kOut[l])) != hipSuccess) return e;
      if ((e = h->da[l].ensure(sizeof(float) * batch * kOut[l])) != hipSuccess) return e;
    }
  }
  h->cap = batch;
  return hipSuccess;
}

// Forward, loss and backward into h->g (gradient of the mean loss).
hipError_t fwd_bwd(eemd_mlp* h, const float* x, const float* y, int batch) {
  hipStream_t st = h->stream;
  const uint16_t* w = (const uint16_t*)h->w16.p;
  float* g = (float*)h->g.p;
  hipError_t e;
  const float* in = x;
  for (int l = 0; l < 4; ++l) {
    float* z = (float*)h->z[l].p;
    if ((e = gemm<float, uint16_t>(in, kIn[l], 0, w + kW[l], kOut[l], 0, z, kOut[l], w + kB[l],
                                   batch, kOut[l], kIn[l], st)) != hipSuccess)
      return e;
    if (l < 3) {
      const int64_t n = (int64_t)batch * kOut[l];
      hipLaunchKernelGGL(gelu_fwd_kernel, dim3(blocks(n, 256)), dim3(256), 0, st, z,
                         (float*)h->a[l].p, n);
      in = (const float*)h->a[l].p;
    }
  }
  hipLaunchKernelGGL(mse_kernel, dim3(1), dim3(256), 0, st, (const float*)h->z[3].p, y,
                     (float*)h->dz[3].p, (float*)h->loss.p, (int64_t)batch * MLP_OUT);
  for (int l = 3; l >= 0; --l) {
    const float* dz = (const float*)h->dz[l].p;
    const float* xin = l == 0 ? x : (const float*)h->a[l - 1].p;
    // dW[in][out] = X^T dZ
    if ((e = gemm<float, float>(xin, kIn[l], 1, dz, kOut[l], 0, g + kW[l], kOut[l], nullptr,
                                kIn[l], kOut[l], batch, st)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(bias_grad_kernel, dim3(blocks(kOut[l], 128)), dim3(128), 0, st, dz,
                       g + kB[l], batch, kOut[l]);
    if (l > 0) {
      // dA = dZ W^T, then through GELU
      if ((e = gemm<float, uint16_t>(dz, kOut[l], 0, w + kW[l], kOut[l], 1, (float*)h->da[l - 1].p,
                                     kIn[l], nullptr, batch, kIn[l], kOut[l], st)) != hipSuccess)
        return e;
      const int64_t n = (int64_t)batch * kIn[l];
      hipLaunchKernelGGL(gelu_bwd_kernel, dim3(blocks(n, 256)), dim3(256), 0, st,
                         (const float*)h->da[l - 1].p, (const float*)h->z[l - 1].p,
                         (float*)h->dz[l - 1].p, n);
    }
  }
  return hipGetLastError();
}

int check_batch(eemd_mlp* h, const float* x, int batch) {
  if (!h) return EEMD_ERR_BAD_ARG;
  if (!x || batch < 16 || batch % 16 != 0) {
    h->err = "batch must be a positive multiple of 16 and x non-NULL";
    return EEMD_ERR_BAD_ARG;
  }
  return EEMD_OK;
}

}  // namespace

extern "C" {

int eemd_mlp_create(int device, uint64_t seed, eemd_mlp** out) {
  if (!out) return EEMD_ERR_BAD_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return EEMD_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev) return EEMD_ERR_BAD_ARG;
  eemd_mlp* h = new eemd_mlp();
  h->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
    (void)hipGetLastError();
    delete h;
    return EEMD_ERR_DEVICE;
  }
  const size_t nb = sizeof(float) * MLP_NPARAM;
  if (h->th.ensure(nb) || h->g.ensure(nb) || h->m.ensure(nb) || h->v.ensure(nb) ||
      h->w16.ensure(sizeof(uint16_t) * MLP_NPARAM) || h->loss.ensure(sizeof(float))) {
    (void)hipGetLastError();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return EEMD_ERR_OOM;
  }
  // LeCun-uniform weights (splitmix64 stream of the seed), zero biases
  std::vector<float> th(MLP_NPARAM, 0.f);
  uint64_t s = seed;
  for (int l = 0; l < 4; ++l) {
    const float a = std::sqrt(3.0f / (float)kIn[l]);
    for (int i = 0; i < kIn[l] * kOut[l]; ++i) {
      s += 0x9E3779B97F4A7C15ull;
      uint64_t z = s;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
      th[kW[l] + i] = (float)((2.0 * u - 1.0) * a);
    }
  }
  *out = h;
  if (eemd_mlp_set(h, th.data()) != EEMD_OK) {
    eemd_mlp_destroy(h);
    *out = nullptr;
    return EEMD_ERR_DEVICE;
  }
  return EEMD_OK;
}

int eemd_mlp_destroy(eemd_mlp* h) {
  if (!h) return EEMD_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamDestroy(h->stream);
  }
  delete h;
  return EEMD_OK;
}

const char* eemd_mlp_last_error(const eemd_mlp* h) { return h ? h->err.c_str() : "null model"; }

int eemd_mlp_set(eemd_mlp* h, const float* params) {
  if (!h || !params) return EEMD_ERR_BAD_ARG;
  MLPCHK(h, hipSetDevice(h->device), "set device");
  const size_t nb = sizeof(float) * MLP_NPARAM;
  MLPCHK(h, hipMemcpyAsync(h->th.p, params, nb, hipMemcpyHostToDevice, h->stream), "upload");
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(MLP_NPARAM, 256)), dim3(256), 0, h->stream,
                     (const float*)h->th.p, (uint16_t*)h->w16.p, MLP_NPARAM);
  MLPCHK(h, hipGetLastError(), "bf16 copy");
  std::vector<float> zero(MLP_NPARAM, 0.f);
  MLPCHK(h, hipMemcpyAsync(h->m.p, zero.data(), nb, hipMemcpyHostToDevice, h->stream), "upload");
  MLPCHK(h, hipMemcpyAsync(h->v.p, zero.data(), nb, hipMemcpyHostToDevice, h->stream), "upload");
  MLPCHK(h, hipStreamSynchronize(h->stream), "upload");
  h->t = 0;
  return EEMD_OK;
}

int eemd_mlp_get(eemd_mlp* h, float* params) {
  if (!h || !params) return EEMD_ERR_BAD_ARG;
  MLPCHK(h, hipSetDevice(h->device), "set device");
  MLPCHK(h, hipMemcpyAsync(params, h->th.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost,
                           h->stream), "download");
  MLPCHK(h, hipStreamSynchronize(h->stream), "download");
  return EEMD_OK;
}

int eemd_mlp_grad(eemd_mlp* h, const float* x, const float* y, int batch, float* grad,
                  float* loss) {
  int rc = check_batch(h, x, batch);
  if (rc) return rc;
  if (!y) return EEMD_ERR_BAD_ARG;
  MLPCHK(h, hipSetDevice(h->device), "set device");
  MLPCHK(h, ensure_batch(h, batch), "activation alloc");
  MLPCHK(h, fwd_bwd(h, x, y, batch), "forward / backward");
  if (grad)
    MLPCHK(h, hipMemcpyAsync(grad, h->g.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost,
                             h->stream), "download");
  if (loss)
    MLPCHK(h, hipMemcpyAsync(loss, h->loss.p, sizeof(float), hipMemcpyDeviceToHost, h->stream),
           "download");
  MLPCHK(h, hipStreamSynchronize(h->stream), "gradient");
  return EEMD_OK;
}

int eemd_mlp_step(eemd_mlp* h, const float* x, const float* y, int batch, float lr, float* loss) {
  int rc = check_batch(h, x, batch);
  if (rc) return rc;
  if (!y || !(lr > 0.f)) return EEMD_ERR_BAD_ARG;
  MLPCHK(h, hipSetDevice(h->device), "set device");
  MLPCHK(h, ensure_batch(h, batch), "activation alloc");
  MLPCHK(h, fwd_bwd(h, x, y, batch), "forward / backward");
  h->t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)h->t), c2 = 1.0f - std::pow(b2, (float)h->t);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks(MLP_NPARAM, 256)), dim3(256), 0, h->stream,
                     (float*)h->th.p, (const float*)h->g.p, (float*)h->m.p, (float*)h->v.p,
                     (uint16_t*)h->w16.p, MLP_NPARAM, lr, b1, b2, 1e-8f, c1, c2);
  MLPCHK(h, hipGetLastError(), "adam");
  if (loss) {
    MLPCHK(h, hipMemcpyAsync(loss, h->loss.p, sizeof(float), hipMemcpyDeviceToHost, h->stream),
           "download");
    MLPCHK(h, hipStreamSynchronize(h->stream), "step");
  }
  return EEMD_OK;
}

int eemd_mlp_infer(eemd_mlp* h, const float* x, int batch, float* y) {
  if (!h || !x || !y || batch < 1) return EEMD_ERR_BAD_ARG;
  MLPCHK(h, hipSetDevice(h->device), "set device");
  hipLaunchKernelGGL(mlp_infer_kernel, dim3(batch), dim3(MLP_HID), 0, h->stream,
                     (const uint16_t*)h->w16.p, x, y, batch);
  MLPCHK(h, hipGetLastError(), "inference");
  MLPCHK(h, hipStreamSynchronize(h->stream), "inference");
  return EEMD_OK;
}

}  // extern "C"
