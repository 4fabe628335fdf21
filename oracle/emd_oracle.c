/*
 * emd_oracle.c -- CPU oracle for the EEMD hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, never the product: only tests/, the smoke entry
 * and the bench's cpu_baseline leg may load it.  The shipped library
 * (libeemd_mi355x.so) never links or calls it.
 *
 * It restates, in plain C with IEEE double arithmetic, no FMA contraction
 * (-ffp-contract=off) and the reference's operation order:
 *   emd_ref.py:31-43   _build_interior_system   -> orc_row()
 *   emd_ref.py:45-48   _rhs                      -> orc_row()
 *   emd_ref.py:50-65   thomas                    -> orc_window_thomas() with one window
 *   emd_gpu_v5_dsbt.py:51-74, 223-237  DSBT      -> orc_window_thomas()/orc_dsbt()
 *   emd_ref.py:67-75   _spline_eval              -> orc_spline_eval()
 *   emd_ref.py:77-103  envelope (natural/mirror) -> orc_envelope()
 *   emd_ref.py:108-156 emd_1d                    -> orc_emd_1d()
 *   emd_ref.py:199-217 eemd (level-wise mean; max_imf is honoured per member,
 *                      as emd_par.py:148, 237 pass it to each one)   -> orc_eemd()
 *   emd_ref.py:141-142 np.sum (8192-element blocks, pairwise inside)  -> orc_pairwise_sum()
 * plus the counter-based noise the GPU path generates in-kernel:
 *   Philox4x32-10 (Salmon et al., SC'11) + Box-Muller                -> orc_philox4x32_10(), orc_normals()
 * and the engine's learned end extension (no reference counterpart): the
 * boundary MLP's exact FP32 inference (orc_mlp_infer), its virtual-extrema
 * rule (orc_learned_virtual) and envelopes with those ends (end mode 2).
 *
 * Deliberate arithmetic choices shared with the HIP FP64 path:
 *   - a**3 (emd_ref.py:75) is numpy's SIMD pow there; here it is the cube
 *     rounded once: hi=a*a, lo=fma(a,a,-hi), cube=fma(hi,a,lo*a).
 *   - divisions are IEEE divisions (never reciprocal multiplies).
 *
 * Extrema trace: for every executed sift the oracle can record the counts of
 * maxima/minima and an order-independent 64-bit hash of their indices
 * (sum of splitmix64(index)), the same digest the GPU trace produces.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_END_NATURAL 0
#define ORC_END_MIRROR 1
#define ORC_END_LEARNED 2

/* ------------------------------------------------------------------ */
/* small helpers                                                        */
/* ------------------------------------------------------------------ */
static inline double orc_cube(double a) {
  double hi = a * a;
  double lo = fma(a, a, -hi);
  return fma(hi, a, lo * a);
}

uint64_t orc_mix64(uint64_t z) {
  /* splitmix64 finaliser (Steele, Lea, Flood 2014) */
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* numpy's pairwise summation of a contiguous float64 block; blocks of <=128
 * use eight interleaved partial sums. */
static double orc_pairwise_block(const double* a, int64_t n) {
  if (n < 8) {
    double r = 0.0;
    for (int64_t i = 0; i < n; i++) r += a[i];
    return r;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; j++) r[j] = a[j];
    int64_t i;
    for (i = 8; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; j++) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += a[i];
    return res;
  }
  int64_t n2 = n / 2;
  n2 -= n2 % 8;
  return orc_pairwise_block(a, n2) + orc_pairwise_block(a + n2, n - n2);
}

/* np.sum of a contiguous float64 vector (emd_ref.py:141-142): the reduction
 * visits the data in buffer-sized blocks of 8192 elements, sums each block
 * pairwise and adds it to the running total. */
double orc_pairwise_sum(const double* a, int64_t n) {
  double acc = 0.0;
  for (int64_t i = 0; i < n; i += 8192) {
    int64_t m = n - i < 8192 ? n - i : 8192;
    acc = acc + orc_pairwise_block(a + i, m);
  }
  return acc;
}

/* ------------------------------------------------------------------ */
/* extrema (emd_ref.py:84-87, 121-122, 131-132)                         */
/* ------------------------------------------------------------------ */
int64_t orc_extrema(const double* h, int64_t N, int want_max, int64_t* k) {
  int64_t p = 0;
  for (int64_t i = 1; i + 1 < N; i++) {
    int hit = want_max ? (h[i] > h[i - 1] && h[i] > h[i + 1])
                       : (h[i] < h[i - 1] && h[i] < h[i + 1]);
    if (hit) {
      if (k) k[p] = i;
      p++;
    }
  }
  return p;
}

/* ------------------------------------------------------------------ */
/* spline system rows (emd_ref.py:31-48)                                */
/* Row q (0..n-1) couples M_{q+1} to M_q, M_{q+2}; H_j = T_{j+1}-T_j.   */
/* ------------------------------------------------------------------ */
static inline void orc_row(const double* T, const double* Y, int64_t n, int64_t q,
                           double* d, double* l, double* u, double* r) {
  double Hq = T[q + 1] - T[q];
  double Hq1 = T[q + 2] - T[q + 1];
  *d = 2.0 * (Hq + Hq1);
  *l = (q >= 1) ? Hq : 0.0;
  *u = (q <= n - 2) ? Hq1 : 0.0;
  *r = 6.0 * ((Y[q + 2] - Y[q + 1]) / Hq1 - (Y[q + 1] - Y[q]) / Hq);
}

/* Thomas over rows [w0, w1] with the couplings outside the window dropped
 * (emd_gpu_v5_dsbt.py:51-74).  With w0=0 and w1=n-1 this is bit-identical
 * to emd_ref.thomas (emd_ref.py:50-65): the first row divides r by d-l*0
 * and the last row's c is 0.  Writes rows [o0, o1] of x. */
void orc_window_thomas(const double* d, const double* l, const double* u,
                       const double* r, int64_t w0, int64_t w1, int64_t o0,
                       int64_t o1, double* cbuf, double* xbuf, double* x) {
  double cprev = 0.0, xprev = 0.0;
  for (int64_t i = w0; i <= w1; i++) {
    double den = d[i] - l[i] * cprev;
    double c = (i == w1) ? 0.0 : u[i] / den;
    double xv = (r[i] - l[i] * xprev) / den;
    cbuf[i - w0] = c;
    xbuf[i - w0] = xv;
    cprev = c;
    xprev = xv;
  }
  double xn = xbuf[w1 - w0];
  if (w1 >= o0 && w1 <= o1) x[w1] = xn;
  for (int64_t i = w1 - 1; i >= w0; i--) {
    double xi = xbuf[i - w0] - cbuf[i - w0] * xn;
    if (i >= o0 && i <= o1) x[i] = xi;
    xn = xi;
  }
}

/* Decay-seamed block Thomas (emd_gpu_v5_dsbt.py:223-237; corrected host
 * mirror audit_eemd/test_block_thomas.py:117-146): owned blocks of B rows,
 * each solved on the clipped window [o0-S, o1+S], owned rows written back.
 * B <= 0 means one window over all rows (exact Thomas). */
void orc_dsbt(const double* d, const double* l, const double* u, const double* r,
              int64_t n, int64_t B, int64_t S, double* x) {
  if (n <= 0) return;
  if (B <= 0 || B >= n) B = n;
  if (S < 0) S = 0;
  int64_t wmax = B + 2 * S;
  if (wmax > n) wmax = n;
  double* cbuf = (double*)malloc(sizeof(double) * (size_t)wmax);
  double* xbuf = (double*)malloc(sizeof(double) * (size_t)wmax);
  int64_t nb = (n + B - 1) / B;
  for (int64_t b = 0; b < nb; b++) {
    int64_t o0 = b * B;
    int64_t o1 = o0 + B;
    if (o1 > n) o1 = n;
    o1 -= 1;
    int64_t w0 = o0 - S;
    if (w0 < 0) w0 = 0;
    int64_t w1 = o1 + S;
    if (w1 > n - 1) w1 = n - 1;
    orc_window_thomas(d, l, u, r, w0, w1, o0, o1, cbuf, xbuf, x);
  }
  free(cbuf);
  free(xbuf);
}

/* Exact Thomas on a caller-built system (emd_ref.py:50-65). */
void orc_thomas(const double* d, const double* l, const double* u, const double* r,
                int64_t n, double* x) {
  orc_dsbt(d, l, u, r, n, 0, 0, x);
}

/* ------------------------------------------------------------------ */
/* spline evaluation (emd_ref.py:67-75) on t = 0..N-1                   */
/* interval idx = clip(searchsorted_left(T, t) - 1, 0, P-2)             */
/* ------------------------------------------------------------------ */
void orc_spline_eval(const double* T, const double* Y, const double* M, int64_t P,
                     int64_t N, double* out) {
  int64_t j = 0; /* number of knots strictly below t */
  for (int64_t i = 0; i < N; i++) {
    double t = (double)i;
    while (j < P && T[j] < t) j++;
    int64_t idx = j - 1;
    if (idx < 0) idx = 0;
    if (idx > P - 2) idx = P - 2;
    double h = T[idx + 1] - T[idx];
    double a = (T[idx + 1] - t) / h;
    double b = (t - T[idx]) / h;
    double ca = orc_cube(a), cb = orc_cube(b);
    out[i] = (a * Y[idx] + b * Y[idx + 1]) +
             ((((ca - a) * M[idx]) + ((cb - b) * M[idx + 1])) * (h * h)) / 6.0;
  }
}

/* ------------------------------------------------------------------ */
/* envelope (emd_ref.py:77-103)                                         */
/* returns p = number of extrema; env[N] written.                        */
/* work must hold 8*(N+2) doubles and (N+2) int64 (kidx).                */
/* ------------------------------------------------------------------ */
typedef struct {
  double* T;
  double* Y;
  double* M;
  double* d;
  double* l;
  double* u;
  double* r;
  double* x;
  int64_t* k;
  double* Tx;  /* learned ends: knots of both envelopes at once */
  double* Yx;
  double* Tn;
  double* Yn;
} orc_work;

static int orc_work_alloc(orc_work* w, int64_t N) {
  size_t m = (size_t)(N + 16);
  w->T = (double*)malloc(sizeof(double) * m);
  w->Y = (double*)malloc(sizeof(double) * m);
  w->M = (double*)malloc(sizeof(double) * m);
  w->d = (double*)malloc(sizeof(double) * m);
  w->l = (double*)malloc(sizeof(double) * m);
  w->u = (double*)malloc(sizeof(double) * m);
  w->r = (double*)malloc(sizeof(double) * m);
  w->x = (double*)malloc(sizeof(double) * m);
  w->k = (int64_t*)malloc(sizeof(int64_t) * m);
  w->Tx = (double*)malloc(sizeof(double) * m);
  w->Yx = (double*)malloc(sizeof(double) * m);
  w->Tn = (double*)malloc(sizeof(double) * m);
  w->Yn = (double*)malloc(sizeof(double) * m);
  return (w->T && w->Y && w->M && w->d && w->l && w->u && w->r && w->x && w->k && w->Tx &&
          w->Yx && w->Tn && w->Yn) ? 0 : -1;
}

static void orc_work_free(orc_work* w) {
  free(w->T); free(w->Y); free(w->M); free(w->d); free(w->l);
  free(w->u); free(w->r); free(w->x); free(w->k);
  free(w->Tx); free(w->Yx); free(w->Tn); free(w->Yn);
}

static uint64_t orc_hash_idx(const int64_t* k, int64_t p) {
  uint64_t s = 0;
  for (int64_t i = 0; i < p; i++) s += orc_mix64((uint64_t)k[i]);
  return s;
}

/* Natural spline through the P extended knots in w->T / w->Y (M = 0 at
 * both outermost knots), evaluated on 0..N-1. */
static void orc_spline_from_knots(orc_work* w, int64_t P, int64_t N, int64_t B, int64_t S,
                                  double* env) {
  int64_t n = P - 2;
  for (int64_t q = 0; q < n; q++) orc_row(w->T, w->Y, n, q, &w->d[q], &w->l[q], &w->u[q], &w->r[q]);
  orc_dsbt(w->d, w->l, w->u, w->r, n, B, S, w->x);
  w->M[0] = 0.0;
  w->M[P - 1] = 0.0;
  for (int64_t q = 0; q < n; q++) w->M[q + 1] = w->x[q];
  orc_spline_eval(w->T, w->Y, w->M, P, N, env);
}

static int64_t orc_envelope_w(const double* h, int64_t N, int want_max, int end,
                              int64_t B, int64_t S, double* env, orc_work* w,
                              uint64_t* hash) {
  int64_t p = orc_extrema(h, N, want_max, w->k);
  if (hash) *hash = orc_hash_idx(w->k, p);
  if (p < 2) {
    double c = (p == 1) ? h[w->k[0]] : 0.0;
    for (int64_t i = 0; i < N; i++) env[i] = c;
    return p;
  }
  int64_t P, off;
  if (end == ORC_END_MIRROR && p >= 3) {
    P = p + 2;
    off = 1;
  } else {
    P = p;
    off = 0;
  }
  for (int64_t j = 0; j < p; j++) {
    w->T[j + off] = (double)w->k[j];
    w->Y[j + off] = h[w->k[j]];
  }
  if (off) {
    double dt0 = w->T[2] - w->T[1];
    double dt1 = w->T[p] - w->T[p - 1];
    w->T[0] = w->T[1] - dt0;
    w->Y[0] = w->Y[2];
    w->T[p + 1] = w->T[p] + dt1;
    w->Y[p + 1] = w->Y[p - 1];
  }
  orc_spline_from_knots(w, P, N, B, S, env);
  return p;
}

/* Public single-envelope entry (tests): end 0 natural, 1 mirror. */
int64_t orc_envelope(const double* h, int64_t N, int want_max, int end, int64_t B,
                     int64_t S, double* env) {
  orc_work w;
  if (orc_work_alloc(&w, N) != 0) return -1;
  int64_t p = orc_envelope_w(h, N, want_max, end, B, S, env, &w, NULL);
  orc_work_free(&w);
  return p;
}

/* Spline pieces exposed for tests: build knots (with the mirror rule),
 * solve, and return T/Y/M of the extended knot set.  Returns P. */
int64_t orc_spline_knots(const double* h, int64_t N, int want_max, int end, int64_t B,
                         int64_t S, double* T, double* Y, double* M) {
  orc_work w;
  if (orc_work_alloc(&w, N) != 0) return -1;
  double* env = (double*)malloc(sizeof(double) * (size_t)N);
  int64_t p = orc_envelope_w(h, N, want_max, end, B, S, env, &w, NULL);
  int64_t P = (p < 2) ? p : ((end == ORC_END_MIRROR && p >= 3) ? p + 2 : p);
  if (p >= 2) {
    memcpy(T, w.T, sizeof(double) * (size_t)P);
    memcpy(Y, w.Y, sizeof(double) * (size_t)P);
    memcpy(M, w.M, sizeof(double) * (size_t)P);
  }
  free(env);
  orc_work_free(&w);
  return P;
}

/* ------------------------------------------------------------------ */
/* Learned end extension (the engine's EEMD_END_LEARNED; no reference    */
/* counterpart).  Model: 64 -> 128 -> 128 -> 128 -> 16, GELU, BF16       */
/* parameters packed per layer as W[in][out] then bias[out].            */
/* Inference arithmetic: each neuron two FP32 fma chains in index order */
/* (first half from the bias, second from 0) added; GELU in tanh form;  */
/* exp by Cody-Waite                                                    */
/* reduction and a degree-7 Taylor polynomial.                          */
/* ------------------------------------------------------------------ */
#define ORC_MLP_IN 64
#define ORC_MLP_HID 128
#define ORC_MLP_OUT 16
#define ORC_CTX 16
#define ORC_PRED 4
#define ORC_W1 0
#define ORC_B1 (ORC_W1 + ORC_MLP_IN * ORC_MLP_HID)
#define ORC_W2 (ORC_B1 + ORC_MLP_HID)
#define ORC_B2 (ORC_W2 + ORC_MLP_HID * ORC_MLP_HID)
#define ORC_W3 (ORC_B2 + ORC_MLP_HID)
#define ORC_B3 (ORC_W3 + ORC_MLP_HID * ORC_MLP_HID)
#define ORC_W4 (ORC_B3 + ORC_MLP_HID)
#define ORC_B4 (ORC_W4 + ORC_MLP_HID * ORC_MLP_OUT)
#define ORC_NPARAM (ORC_B4 + ORC_MLP_OUT)

static const uint16_t* orc_model_w = NULL;

/* Model used by end mode 2 (BF16 parameters; NULL: mirror everywhere). */
void orc_set_model(const uint16_t* w) { orc_model_w = w; }
int orc_mlp_nparam(void) { return ORC_NPARAM; }

float orc_bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

uint16_t orc_f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

float orc_mlp_exp(float x) {
  x = fminf(fmaxf(x, -87.0f), 88.0f);
  const float k = rintf(x * 1.44269504088896341f);
  float r = fmaf(-k, 0.693145751953125f, x);
  r = fmaf(-k, 1.42860682030941723e-6f, r);
  float p = 1.0f / 5040.0f;
  p = fmaf(p, r, 1.0f / 720.0f);
  p = fmaf(p, r, 1.0f / 120.0f);
  p = fmaf(p, r, 1.0f / 24.0f);
  p = fmaf(p, r, 1.0f / 6.0f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)k);
}

float orc_mlp_gelu(float x) {
  const float x3 = (x * x) * x;
  const float inner = fmaf(0.044715f, x3, x);
  float u = 0.7978845608028654f * inner;
  u = fminf(fmaxf(u, -15.0f), 15.0f);
  const float e = orc_mlp_exp(u + u);
  const float t = 1.0f - 2.0f / (1.0f + e);
  return (0.5f * x) * (1.0f + t);
}

/* bias + sum w[i*ld] in[i]: fma chains lo (i < n/2, from the bias) and hi
 * (i >= n/2, from 0) in index order, then lo + hi. */
static float orc_neuron(const uint16_t* w, int ld, float bias, const float* in, int n) {
  float lo = bias, hi = 0.0f;
  for (int i = 0; i < n / 2; i++) lo = fmaf(orc_bf16_to_f32(w[(int64_t)i * ld]), in[i], lo);
  for (int i = n / 2; i < n; i++) hi = fmaf(orc_bf16_to_f32(w[(int64_t)i * ld]), in[i], hi);
  return lo + hi;
}

void orc_mlp_infer(const uint16_t* w, const float* x, float* y) {
  float h0[ORC_MLP_HID], h1[ORC_MLP_HID];
  for (int j = 0; j < ORC_MLP_HID; j++)
    h0[j] = orc_mlp_gelu(orc_neuron(w + ORC_W1 + j, ORC_MLP_HID, orc_bf16_to_f32(w[ORC_B1 + j]), x,
                                    ORC_MLP_IN));
  for (int j = 0; j < ORC_MLP_HID; j++)
    h1[j] = orc_mlp_gelu(orc_neuron(w + ORC_W2 + j, ORC_MLP_HID, orc_bf16_to_f32(w[ORC_B2 + j]),
                                    h0, ORC_MLP_HID));
  for (int j = 0; j < ORC_MLP_HID; j++)
    h0[j] = orc_mlp_gelu(orc_neuron(w + ORC_W3 + j, ORC_MLP_HID, orc_bf16_to_f32(w[ORC_B3 + j]),
                                    h1, ORC_MLP_HID));
  for (int j = 0; j < ORC_MLP_OUT; j++)
    y[j] = orc_neuron(w + ORC_W4 + j, ORC_MLP_OUT, orc_bf16_to_f32(w[ORC_B4 + j]), h0, ORC_MLP_HID);
}

/* In-place sum of 32 floats by the xor-butterfly tree: for partner offsets
 * 16, 8, 4, 2, 1, every slot s becomes v[s] + v[s ^ k]; v[0] ends as the sum. */
static void orc_butterfly32(float* v) {
  for (int k = 16; k >= 1; k >>= 1) {
    float nv[32];
    for (int s = 0; s < 32; s++) nv[s] = v[s] + v[s ^ k];
    memcpy(v, nv, sizeof(nv));
  }
}

/* Model input of one end (0: left, 1: right) from the knots of both
 * envelopes (time order), and the scales the output is decoded with. */
void orc_learned_features(const double* Tx, const double* Yx, int64_t px, const double* Tn,
                          const double* Yn, int64_t pn, int64_t N, int end, float* x,
                          double* ls_out, float* scale) {
  const double* Ts[2] = {Tx, Tn};
  const double* Ys[2] = {Yx, Yn};
  const int64_t ps[2] = {px, pn};
  double ls[2][ORC_CTX], ly[2][ORC_CTX];
  for (int ev = 0; ev < 2; ev++)
    for (int i = 0; i < ORC_CTX; i++) {
      const int64_t g = end ? ps[ev] - 1 - i : i;
      const double t = Ts[ev][g];
      ls[ev][i] = end ? (double)(N - 1) - t : t;
      ly[ev][i] = Ys[ev][g];
    }
  /* sums over the 32 [env][i] slots: xor-butterfly tree, partner offsets
   * 16, 8, 4, 2, 1 (slot s = env * 16 + i) */
  float gv[32], yv[32], vv[32], gap[32];
  for (int ev = 0; ev < 2; ev++)
    for (int i = 0; i < ORC_CTX; i++) {
      gap[ev * 16 + i] = (float)(i == 0 ? ls[ev][0] : ls[ev][i] - ls[ev][i - 1]);
      gv[ev * 16 + i] = (i == 0) ? 0.0f : gap[ev * 16 + i];
      yv[ev * 16 + i] = (float)ly[ev][i];
    }
  orc_butterfly32(gv);
  orc_butterfly32(yv);
  const float G = gv[0] / (float)(2 * (ORC_CTX - 1));
  const float mu = yv[0] / (float)(2 * ORC_CTX);
  for (int s = 0; s < 32; s++) {
    const float d = (float)ly[s >> 4][s & 15] - mu;
    vv[s] = d * d;
  }
  orc_butterfly32(vv);
  const float sd = fmaxf(sqrtf(vv[0] / (float)(2 * ORC_CTX)), 1e-30f);
  for (int ev = 0; ev < 2; ev++)
    for (int i = 0; i < ORC_CTX; i++) {
      x[ev * 32 + i] = gap[ev * 16 + i] / G - 1.0f;
      x[ev * 32 + 16 + i] = ((float)ly[ev][i] - mu) / sd;
    }
  if (ls_out) {
    ls_out[0] = ls[0][0];
    ls_out[1] = ls[1][0];
  }
  scale[0] = G;
  scale[1] = mu;
  scale[2] = sd;
}

/* Virtual extrema of both envelopes at both ends: vt / vy [env][end][k],
 * k = 0 nearest to the signal.  Gaps: first from the end sample,
 * max(G exp(o) - 1, 0), kept >= 2 from the nearest extremum; then
 * max(G exp(o), 2) apart (o clamped to [-4, 4]); values mu + sd o. */
void orc_learned_virtual(const double* Tx, const double* Yx, int64_t px, const double* Tn,
                         const double* Yn, int64_t pn, int64_t N, const uint16_t* w,
                         double* vt, double* vy) {
  for (int end = 0; end < 2; end++) {
    float x[ORC_MLP_IN], o[ORC_MLP_OUT], sc[3];
    double ls0[2];
    orc_learned_features(Tx, Yx, px, Tn, Yn, pn, N, end, x, ls0, sc);
    orc_mlp_infer(w, x, o);
    const float G = sc[0], mu = sc[1], sd = sc[2];
    for (int ev = 0; ev < 2; ev++) {
      const float* oe = o + ev * 8;
      const double dir = end ? 1.0 : -1.0;
      const double e_end = end ? (double)(N - 1) : 0.0;
      const double t_near = e_end - dir * ls0[ev];
      const float d0 = fmaxf(G * orc_mlp_exp(fminf(fmaxf(oe[0], -4.0f), 4.0f)) - 1.0f, 0.0f);
      double q = e_end + dir * (double)d0;
      if (dir * (q - t_near) < 2.0) q = t_near + dir * 2.0;
      double* pt = vt + (ev * 2 + end) * ORC_PRED;
      double* py = vy + (ev * 2 + end) * ORC_PRED;
      pt[0] = q;
      for (int k = 1; k < ORC_PRED; k++) {
        const float gk = fmaxf(G * orc_mlp_exp(fminf(fmaxf(oe[k], -4.0f), 4.0f)), 2.0f);
        q = q + dir * (double)gk;
        pt[k] = q;
      }
      for (int k = 0; k < ORC_PRED; k++) py[k] = (double)(mu + sd * oe[4 + k]);
    }
  }
}

/* Both envelopes of h with learned ends (mirror when an envelope has fewer
 * than ORC_CTX extrema or no model is set). */
static void orc_envelopes_learned(const double* h, int64_t N, int64_t B, int64_t S, double* eu,
                                  double* el, orc_work* w, uint64_t* hmx, uint64_t* hmn) {
  const int64_t px = orc_extrema(h, N, 1, w->k);
  if (hmx) *hmx = orc_hash_idx(w->k, px);
  for (int64_t j = 0; j < px; j++) { w->Tx[j] = (double)w->k[j]; w->Yx[j] = h[w->k[j]]; }
  const int64_t pn = orc_extrema(h, N, 0, w->k);
  if (hmn) *hmn = orc_hash_idx(w->k, pn);
  for (int64_t j = 0; j < pn; j++) { w->Tn[j] = (double)w->k[j]; w->Yn[j] = h[w->k[j]]; }
  if (!orc_model_w || px < ORC_CTX || pn < ORC_CTX) {
    orc_envelope_w(h, N, 1, ORC_END_MIRROR, B, S, eu, w, NULL);
    orc_envelope_w(h, N, 0, ORC_END_MIRROR, B, S, el, w, NULL);
    return;
  }
  double vt[2 * 2 * ORC_PRED], vy[2 * 2 * ORC_PRED];
  orc_learned_virtual(w->Tx, w->Yx, px, w->Tn, w->Yn, pn, N, orc_model_w, vt, vy);
  for (int ev = 0; ev < 2; ev++) {
    const double* Tk = ev ? w->Tn : w->Tx;
    const double* Yk = ev ? w->Yn : w->Yx;
    const int64_t p = ev ? pn : px;
    const double* lt = vt + (ev * 2 + 0) * ORC_PRED;
    const double* ly = vy + (ev * 2 + 0) * ORC_PRED;
    const double* rt = vt + (ev * 2 + 1) * ORC_PRED;
    const double* ry = vy + (ev * 2 + 1) * ORC_PRED;
    for (int k = 0; k < ORC_PRED; k++) {
      w->T[k] = lt[ORC_PRED - 1 - k];
      w->Y[k] = ly[ORC_PRED - 1 - k];
    }
    for (int64_t j = 0; j < p; j++) {
      w->T[ORC_PRED + j] = Tk[j];
      w->Y[ORC_PRED + j] = Yk[j];
    }
    for (int k = 0; k < ORC_PRED; k++) {
      w->T[ORC_PRED + p + k] = rt[k];
      w->Y[ORC_PRED + p + k] = ry[k];
    }
    orc_spline_from_knots(w, p + 2 * ORC_PRED, N, B, S, ev ? el : eu);
  }
}

/* Array forms for tests. */
void orc_mlp_exp_n(const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; i++) y[i] = orc_mlp_exp(x[i]);
}
void orc_mlp_gelu_n(const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; i++) y[i] = orc_mlp_gelu(x[i]);
}
void orc_mlp_infer_n(const uint16_t* w, const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; i++) orc_mlp_infer(w, x + i * ORC_MLP_IN, y + i * ORC_MLP_OUT);
}
void orc_f32_to_bf16_n(const float* x, uint16_t* y, int64_t n) {
  for (int64_t i = 0; i < n; i++) y[i] = orc_f32_to_bf16(x[i]);
}

/* Public: both envelopes with learned ends (tests). */
int orc_envelopes_learned_pub(const double* h, int64_t N, int64_t B, int64_t S, double* eu,
                              double* el) {
  orc_work w;
  if (orc_work_alloc(&w, N) != 0) return -1;
  orc_envelopes_learned(h, N, B, S, eu, el, &w, NULL, NULL);
  orc_work_free(&w);
  return 0;
}

/* ------------------------------------------------------------------ */
/* single-signal EMD (emd_ref.py:108-156)                               */
/* imfs: [max_imf_cap, N] output; per_mode_sifts[max_imf_cap].           */
/* trace (optional): per executed sift 4 uint64 {nmax, nmin, hmax, hmin} */
/*   laid out [max_imf_cap][max_sifts][4].                               */
/* max_imf = -1 -> decompose until extrema run out (capped at cap).      */
/* Returns n_modes (or -1 on allocation failure).                        */
/* ------------------------------------------------------------------ */
int orc_emd_1d(const double* x, int64_t N, double tau, int max_sifts, int max_imf,
               int min_extrema, int end, int64_t B, int64_t S, int max_imf_cap,
               double* imfs, double* resid_out, int* per_mode_sifts,
               int* total_sifts_out, uint64_t* trace) {
  orc_work w;
  if (orc_work_alloc(&w, N) != 0) return -1;
  double* resid = (double*)malloc(sizeof(double) * (size_t)N);
  double* h = (double*)malloc(sizeof(double) * (size_t)N);
  double* hn = (double*)malloc(sizeof(double) * (size_t)N);
  double* eu = (double*)malloc(sizeof(double) * (size_t)N);
  double* el = (double*)malloc(sizeof(double) * (size_t)N);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)N);
  memcpy(resid, x, sizeof(double) * (size_t)N);
  int n_modes = 0, total = 0;
  for (;;) {
    int64_t nmax = orc_extrema(resid, N, 1, NULL);
    int64_t nmin = orc_extrema(resid, N, 0, NULL);
    if (nmax + nmin < min_extrema) break;
    if (max_imf != -1 && n_modes >= max_imf) break;
    if (n_modes >= max_imf_cap) break;
    memcpy(h, resid, sizeof(double) * (size_t)N);
    int have_prev = 0;
    int n_sift = 0;
    for (int s = 0; s < max_sifts; s++) {
      int64_t cmax = orc_extrema(h, N, 1, NULL);
      int64_t cmin = orc_extrema(h, N, 0, NULL);
      if (cmax < 1 || cmin < 1) break;
      uint64_t hmx = 0, hmn = 0;
      if (end == ORC_END_LEARNED) {
        orc_envelopes_learned(h, N, B, S, eu, el, &w, &hmx, &hmn);
      } else {
        orc_envelope_w(h, N, 1, end, B, S, eu, &w, &hmx);
        orc_envelope_w(h, N, 0, end, B, S, el, &w, &hmn);
      }
      if (trace) {
        uint64_t* tr = trace + ((size_t)n_modes * (size_t)max_sifts + (size_t)s) * 4;
        tr[0] = (uint64_t)cmax; tr[1] = (uint64_t)cmin; tr[2] = hmx; tr[3] = hmn;
      }
      for (int64_t i = 0; i < N; i++) {
        double m = 0.5 * (eu[i] + el[i]);
        hn[i] = h[i] - m;
      }
      n_sift++;
      if (have_prev) {
        /* h_prev is the h this sift started from (emd_ref.py:140-147) */
        for (int64_t i = 0; i < N; i++) { double dd = h[i] - hn[i]; tmp[i] = dd * dd; }
        double num = orc_pairwise_sum(tmp, N);
        for (int64_t i = 0; i < N; i++) tmp[i] = h[i] * h[i];
        double den = orc_pairwise_sum(tmp, N);
        double sd = (den > 0) ? num / den : 0.0;
        if (sd < tau) {
          memcpy(h, hn, sizeof(double) * (size_t)N);
          break;
        }
      }
      have_prev = 1;
      memcpy(h, hn, sizeof(double) * (size_t)N);
    }
    memcpy(imfs + (size_t)n_modes * (size_t)N, h, sizeof(double) * (size_t)N);
    per_mode_sifts[n_modes] = n_sift;
    total += n_sift;
    n_modes++;
    for (int64_t i = 0; i < N; i++) resid[i] = resid[i] - h[i];
  }
  if (resid_out) memcpy(resid_out, resid, sizeof(double) * (size_t)N);
  if (total_sifts_out) *total_sifts_out = total;
  free(resid); free(h); free(hn); free(eu); free(el); free(tmp);
  orc_work_free(&w);
  return n_modes;
}

/* ------------------------------------------------------------------ */
/* EEMD (emd_ref.py:199-217 with max_imf honoured per member):           */
/* member e decomposes x + z[e]*sigma; level-wise sequential mean.       */
/* z: [E, N] standard normals.  avg: [max_imf_cap, N]; resid: [N].       */
/* nmode[E], nsift[E]; per_mode[E*max_imf_cap] optional;                 */
/* member_imfs [E, cap, N] optional; trace [E, cap, max_sifts, 4] opt.   */
/* nthreads: OpenMP threads over members (<=0: runtime default).         */
/* Returns n_modes of the mean (levels with count > 0).                  */
/* ------------------------------------------------------------------ */
int orc_eemd(const double* x, int64_t N, int E, const double* z, double sigma,
             double tau, int max_sifts, int max_imf, int min_extrema, int end,
             int64_t B, int64_t S, int max_imf_cap, int nthreads, double* avg,
             double* resid, int* nmode, int* nsift, int* per_mode,
             double* member_imfs, uint64_t* trace) {
  size_t per = (size_t)max_imf_cap * (size_t)N;
  double* imfs = member_imfs;
  int own = 0;
  if (!imfs) {
    imfs = (double*)malloc(sizeof(double) * per * (size_t)E);
    own = 1;
    if (!imfs) return -1;
  }
  int* pm = per_mode;
  int own_pm = 0;
  if (!pm) {
    pm = (int*)malloc(sizeof(int) * (size_t)E * (size_t)max_imf_cap);
    own_pm = 1;
  }
  int fail = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int e = 0; e < E; e++) {
    double* xe = (double*)malloc(sizeof(double) * (size_t)N);
    for (int64_t i = 0; i < N; i++) xe[i] = x[i] + z[(size_t)e * (size_t)N + (size_t)i] * sigma;
    int ts = 0;
    uint64_t* tr = trace ? trace + (size_t)e * (size_t)max_imf_cap * (size_t)max_sifts * 4 : NULL;
    int nm = orc_emd_1d(xe, N, tau, max_sifts, max_imf, min_extrema, end, B, S, max_imf_cap,
                        imfs + per * (size_t)e, NULL, pm + (size_t)e * (size_t)max_imf_cap,
                        &ts, tr);
    if (nm < 0) fail = 1;
    nmode[e] = nm;
    nsift[e] = ts;
    free(xe);
  }
  if (fail) {
    if (own) free(imfs);
    if (own_pm) free(pm);
    return -1;
  }
  int maxm = 0;
  for (int e = 0; e < E; e++) if (nmode[e] > maxm) maxm = nmode[e];
  for (int m = 0; m < max_imf_cap; m++) {
    double* a = avg + (size_t)m * (size_t)N;
    for (int64_t i = 0; i < N; i++) a[i] = 0.0;
    int cnt = 0;
    for (int e = 0; e < E; e++) {
      if (nmode[e] > m) {
        const double* im = imfs + per * (size_t)e + (size_t)m * (size_t)N;
        for (int64_t i = 0; i < N; i++) a[i] += im[i];
        cnt++;
      }
    }
    double c = (cnt > 0) ? (double)cnt : 1.0;
    for (int64_t i = 0; i < N; i++) a[i] = a[i] / c;
  }
  /* resid = x - avg[:n_modes].sum(axis=0); numpy reduces axis 0 row by row */
  for (int64_t i = 0; i < N; i++) {
    double s = 0.0;
    if (maxm > 0) {
      s = avg[i];
      for (int m = 1; m < maxm; m++) s += avg[(size_t)m * (size_t)N + (size_t)i];
    }
    resid[i] = x[i] - s;
  }
  if (own) free(imfs);
  if (own_pm) free(pm);
  return maxm;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller     */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int round = 0; round < 10; round++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Standard normals for (signal, member), samples 0..N-1.  Sample pair j
 * (samples 2j, 2j+1) uses counter {j, member, signal, 0x45454D44} and key
 * {seed lo, seed hi}; u1 = (top 53 bits of w0:w1 + 1) * 2^-53 in (0,1],
 * u2 = top 53 bits of w2:w3 * 2^-53 in [0,1);
 * z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = sqrt(-2 ln u1) sin(2 pi u2). */
void orc_normals(uint64_t seed, uint32_t signal, uint32_t member, int64_t N, double* out) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t j = 0; 2 * j < N; j++) {
    uint32_t ctr[4] = {(uint32_t)j, member, signal, 0x45454D44u};
    uint32_t w[4];
    orc_philox4x32_10(ctr, key, w);
    uint64_t a = (((uint64_t)w[0] << 32) | w[1]) >> 11;
    uint64_t b = (((uint64_t)w[2] << 32) | w[3]) >> 11;
    double u1 = ((double)a + 1.0) * 0x1.0p-53;
    double u2 = (double)b * 0x1.0p-53;
    double rad = sqrt(-2.0 * log(u1));
    double ang = two_pi * u2;
    out[2 * j] = rad * cos(ang);
    if (2 * j + 1 < N) out[2 * j + 1] = rad * sin(ang);
  }
}
