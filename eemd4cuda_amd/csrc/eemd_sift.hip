// Persistent EEMD sift engine for gfx950 (CDNA4, wave64).
//
// One workgroup pops tasks (signal, member range) from a global queue and runs
// the complete EMD of each member: noise add -> [extrema -> spline solve ->
// envelope evaluation + mean subtraction + stopping test]* -> IMF / residual
// update -> ensemble accumulation.  Reference behaviour followed:
//   extrema            emd_ref.py:84-87, 121-122, 131-132 (strict, time ordered)
//   mirror extension   emd_ref.py:92-95
//   spline system      emd_ref.py:31-48
//   Thomas / DSBT      emd_ref.py:50-65, emd_gpu_v5_dsbt.py:51-74, 223-237
//   spline evaluation  emd_ref.py:67-75 (searchsorted-left intervals)
//   sift / IMF loop    emd_ref.py:108-156
//   ensemble mean      emd_ref.py:209-217, emd_par.py:241-250
//
// Data layout (per workgroup scratch, global memory):
//   h, r        working h and residual, T[npad]
//   bits        1 bit per sample per envelope: bit b <=> sample b-1 is an
//               extremum (the shift lets a chunk decide its bits from values
//               it has already produced); two sets (current / next sift)
//   kt, ky      compacted knots per (set, envelope), stored per segment region
//               (the wave that owns a signal segment writes its knots to its
//               own region, so no barrier is needed inside a pass)
//   km          spline second derivatives of the current knots
//   wc, wx, ws  forward sweep of the window Thomas solves
//
// FP64 mode follows the oracle's arithmetic exactly (no contraction: the file
// is compiled with -ffp-contract=off; IEEE divisions; the cube rounded once).
// FP32 mode stores samples and knots in float, evaluates envelopes in float
// with FMAs, and solves the spline systems in double.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "eemd_common.h"

namespace eemd {

template <typename T>
struct Knot {
  T t, y, m;
  int code;  // >= 0: real knot slot (relative to the envelope region); < 0: virtual
};

template <typename T, int NW>
struct Shared {
  int32_t cnt[2][2][NW];
  int32_t beg[2][2][NW];
  int32_t pre[2][2][NW + 1];
  unsigned long long hsh[2][2];
  int32_t p[2], nl[2], nr[2], P[2];
  T cval[2];
  T vpos[2][2 * MAX_VIRT], vy[2][2 * MAX_VIRT], vm[2][2 * MAX_VIRT];
  double red[2][NW];
  int32_t acc_cnt[MAX_IMF_CAP];
  int32_t pm[MAX_IMF_CAP];
  int32_t task;
  double sd;
  T stage[NW * RPL * WAVE];  // per-wave envelope staging rows
};

// ------------------------------------------------------------------------
// wave primitives
// ------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Exclusive prefix and total over the wave of small counts (0..15), built
// from four ballots and masked popcounts (no LDS traffic).
__device__ __forceinline__ int wave_excl_small(int c, int& total) {
  const int ln = lane_id();
  const unsigned long long lt = (ln == 0) ? 0ull : (~0ull >> (64 - ln));
  int ex = 0, tot = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const unsigned long long m = __ballot((c >> b) & 1);
    ex += __popcll(m & lt) << b;
    tot += __popcll(m) << b;
  }
  total = tot;
  return ex;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// a**3 rounded once (hi + lo = a*a exactly; one fma adds lo*a).
__device__ __forceinline__ double cube_once(double a) {
  const double hi = a * a;
  const double lo = __builtin_fma(a, a, -hi);
  return __builtin_fma(hi, a, lo * a);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c[0];
    const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2];
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Standard normal pair for sample pair j of (signal, member).
__device__ __forceinline__ void normal_pair(uint64_t seed, uint32_t sig, uint32_t mem,
                                            uint32_t j, double& z0, double& z1) {
  uint32_t c[4] = {j, mem, sig, 0x45454D44u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const unsigned long long a = ((((unsigned long long)c[0]) << 32) | c[1]) >> 11;
  const unsigned long long b = ((((unsigned long long)c[2]) << 32) | c[3]) >> 11;
  const double u1 = ((double)a + 1.0) * 0x1.0p-53;
  const double u2 = (double)b * 0x1.0p-53;
  const double rad = sqrt(-2.0 * log(u1));
  const double ang = 6.283185307179586476925286766559 * u2;
  double s, co;
  sincos(ang, &s, &co);
  z0 = rad * co;
  z1 = rad * s;
}

template <typename T>
__device__ __forceinline__ void load8(const T* p, T v[8]) {
  if constexpr (std::is_same<T, float>::value) {
    const float4 a = *reinterpret_cast<const float4*>(p);
    const float4 b = *reinterpret_cast<const float4*>(p + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double2 a = *reinterpret_cast<const double2*>(p + 2 * q);
      v[2 * q] = a.x;
      v[2 * q + 1] = a.y;
    }
  }
}

template <typename T>
__device__ __forceinline__ void store8(T* p, const T v[8]) {
  if constexpr (std::is_same<T, float>::value) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float4*>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      *reinterpret_cast<double2*>(p + 2 * q) = make_double2(v[2 * q], v[2 * q + 1]);
  }
}

// ------------------------------------------------------------------------
// per-member context
// ------------------------------------------------------------------------
template <typename T>
struct Mem {
  int64_t n, sl, cap;
  T* h;
  T* r;
  uint8_t* bits;   // [2 sets][2 env][npad/8]
  int64_t bstride; // npad/8
  int32_t* kt;     // [2][2][NW*cap]
  T* ky;
  T* km;           // [2 env][NW*cap]
  int64_t rstride; // NW*cap
  double* wc;
  double* wx;
  int32_t* ws;
  int64_t cols;
};

template <typename T, int NW>
__device__ __forceinline__ Knot<T> fetch_knot(const Shared<T, NW>& S, const Mem<T>& M, int set,
                                              int env, int e, int& whint) {
  Knot<T> k;
  const int nl = S.nl[env];
  const int p = S.p[env];
  if (e < nl) {
    k.t = S.vpos[env][e];
    k.y = S.vy[env][e];
    k.m = S.vm[env][e];
    k.code = -1 - e;
    return k;
  }
  const int g = e - nl;
  if (g >= p) {
    const int j = g - p;
    k.t = S.vpos[env][MAX_VIRT + j];
    k.y = S.vy[env][MAX_VIRT + j];
    k.m = S.vm[env][MAX_VIRT + j];
    k.code = -1 - MAX_VIRT - j;
    return k;
  }
  const int32_t* pre = S.pre[set][env];
  int w = whint;
#pragma unroll 1
  while (w > 0 && g < pre[w]) --w;
#pragma unroll 1
  while (w < NW - 1 && g >= pre[w + 1]) ++w;
  whint = w;
  const int64_t slot = (int64_t)w * M.cap + S.beg[set][env][w] + (g - pre[w]);
  const int64_t base = (int64_t)(set * 2 + env) * M.rstride;
  k.t = (T)M.kt[base + slot];
  k.y = M.ky[base + slot];
  k.m = M.km[(int64_t)env * M.rstride + slot];
  k.code = (int)slot;
  return k;
}

// Interval (extended-knot index of its left knot) of sample i given
// cl = number of real knots at positions < i: clip(count_ext(<i) - 1, 0, P-2),
// i.e. searchsorted-left minus one (emd_ref.py:70).
template <typename T, int NW>
__device__ __forceinline__ int ext_interval(const Shared<T, NW>& S, int env, int cl, int i) {
  const int nl = S.nl[env], nr = S.nr[env], p = S.p[env];
  int ce = cl + nl;
  if (cl == 0 && nl > 0) {
    ce = 0;
    for (int j = 0; j < nl; ++j) ce += (S.vpos[env][j] < (T)i) ? 1 : 0;
  }
  if (cl == p && nr > 0) {
    for (int j = 0; j < nr; ++j) ce += (S.vpos[env][MAX_VIRT + j] < (T)i) ? 1 : 0;
  }
  int e = ce - 1;
  const int hi = S.P[env] - 2;
  e = e < 0 ? 0 : e;
  e = e > hi ? hi : e;
  return e;
}

// Cubic spline value on [k0, k1] at t (emd_ref.py:71-75).
__device__ __forceinline__ double spline_val(const Knot<double>& k0, const Knot<double>& k1,
                                             double t) {
  const double hh = k1.t - k0.t;
  const double a = (k1.t - t) / hh;
  const double b = (t - k0.t) / hh;
  const double ca = cube_once(a), cb = cube_once(b);
  return (a * k0.y + b * k1.y) + ((((ca - a) * k0.m) + ((cb - b) * k1.m)) * (hh * hh)) / 6.0;
}

template <typename T>
struct Cursor {
  Knot<T> k0, k1;
  int e;
  int wh;
  T ih, h26;  // float path: 1/h and h*h/6 of the current interval
};

template <typename T>
__device__ __forceinline__ void cursor_prep(Cursor<T>& c) {
  if constexpr (std::is_same<T, float>::value) {
    const float hh = c.k1.t - c.k0.t;
    c.ih = 1.0f / hh;
    c.h26 = hh * hh * (1.0f / 6.0f);
  }
}

template <typename T>
__device__ __forceinline__ T cursor_eval(const Cursor<T>& c, int i) {
  if constexpr (std::is_same<T, double>::value) {
    return spline_val(c.k0, c.k1, (double)i);
  } else {
    const float t = (float)i;
    const float a = (c.k1.t - t) * c.ih;
    const float b = (t - c.k0.t) * c.ih;
    const float ca = __builtin_fmaf(a, a, -1.0f) * a;
    const float cb = __builtin_fmaf(b, b, -1.0f) * b;
    const float lin = __builtin_fmaf(a, c.k0.y, b * c.k1.y);
    return __builtin_fmaf(__builtin_fmaf(ca, c.k0.m, cb * c.k1.m), c.h26, lin);
  }
}

// One envelope's values at samples i0..i0+7, written (ev 0) or added (ev 1)
// into the wave's LDS staging rows stage[k*64 + lane].  base = number of
// real knots of set `cur` with bit index < i0; byte = the lane's 8 bits of
// the current extrema bitmap.  The sample loop is rolled: the interval
// cursor advances one knot at a time (emd_ref.py:70 searchsorted-left).
template <typename T, int NW>
__device__ __forceinline__ void envelope_chunk(const Shared<T, NW>& S, const Mem<T>& M, int cur,
                                               int ev, int w, int base, int byte, int i0,
                                               T* stage) {
  const int ln = lane_id();
  if (S.p[ev] < 2) {
    const T c = S.cval[ev];
#pragma unroll 1
    for (int k = 0; k < RPL; ++k) stage[k * WAVE + ln] = ev == 0 ? c : stage[k * WAVE + ln] + c;
    return;
  }
  Cursor<T> cs;
  cs.wh = w;
  cs.e = -2;
#pragma unroll 1
  for (int k = 0; k < RPL; ++k) {
    const int cl = base + __popc(byte & ((2u << k) - 1u));
    const int e = ext_interval(S, ev, cl, i0 + k);
    if (e != cs.e) {
      if (e != cs.e + 1) cs.k1 = fetch_knot(S, M, cur, ev, e, cs.wh);
      cs.k0 = cs.k1;
      cs.k1 = fetch_knot(S, M, cur, ev, e + 1, cs.wh);
      cs.e = e;
      cursor_prep(cs);
    }
    const T ek = cursor_eval(cs, i0 + k);
    stage[k * WAVE + ln] = ev == 0 ? ek : stage[k * WAVE + ln] + ek;
  }
}

// ------------------------------------------------------------------------
// passes
// ------------------------------------------------------------------------
enum PassMode { PASS_INIT = 0, PASS_SIFT = 1, PASS_LAST = 2, PASS_FINAL = 3 };

struct PassIO {
  const double* x;      // INIT: signal
  const double* z;      // INIT: explicit normals row (or null)
  double sigma;
  uint64_t seed;
  uint32_t sig, mem;
  int src_is_r;         // SIFT/LAST/FINAL: h is read from r
  int want_sd;          // SIFT: accumulate the SD sums
  void* imf;            // LAST/FINAL: member IMF row (T) or null
  double* acc;          // LAST/FINAL: task accumulator row or null
  int acc_first;        // first contribution to this level in the task
  int trace;            // accumulate knot digests
};

// One wave streams its segment: produce new values, write them, detect the
// extrema of the new values and compact them into set `nxt`.
template <typename T, int NW, int MODE>
__device__ void wave_pass(Shared<T, NW>& S, const Mem<T>& M, const PassIO& io, int cur, int nxt) {
  const int ln = lane_id();
  const int w = (int)(threadIdx.x / WAVE);
  const int64_t n = M.n;
  const int64_t seg_lo = (int64_t)w * M.sl;
  int64_t seg_hi = seg_lo + M.sl;
  if (seg_hi > n) seg_hi = n;

  int run[2] = {0, 0};      // knots of set nxt written in this segment
  int crun[2] = {0, 0};     // knots of set cur passed in this segment (eval)
  unsigned long long hs[2] = {0ull, 0ull};
  double num = 0.0, den = 0.0;
  T carry1 = (T)0, carry2 = (T)0;

  T* const dst = (MODE == PASS_SIFT) ? M.h : M.r;
  const T* const src = io.src_is_r ? M.r : M.h;
  int32_t* const ktn[2] = {M.kt + (int64_t)(nxt * 2 + 0) * M.rstride + (int64_t)w * M.cap,
                           M.kt + (int64_t)(nxt * 2 + 1) * M.rstride + (int64_t)w * M.cap};
  T* const kyn[2] = {M.ky + (int64_t)(nxt * 2 + 0) * M.rstride + (int64_t)w * M.cap,
                     M.ky + (int64_t)(nxt * 2 + 1) * M.rstride + (int64_t)w * M.cap};
  uint8_t* const bn[2] = {M.bits + (int64_t)(nxt * 2 + 0) * M.bstride,
                          M.bits + (int64_t)(nxt * 2 + 1) * M.bstride};
  const uint8_t* const bc[2] = {M.bits + (int64_t)(cur * 2 + 0) * M.bstride,
                                M.bits + (int64_t)(cur * 2 + 1) * M.bstride};
  const int kcap = (int)(M.cap - SEG_FRONT);

  for (int64_t cb = seg_lo; cb < seg_hi; cb += CHUNK) {
    const int64_t i0 = cb + (int64_t)ln * RPL;
    T v[RPL];

    if constexpr (MODE == PASS_INIT) {
      double zz[RPL];
      if (io.z) {
#pragma unroll
        for (int k = 0; k < RPL; ++k) zz[k] = (i0 + k < n) ? io.z[i0 + k] : 0.0;
      } else {
#pragma unroll
        for (int q = 0; q < RPL / 2; ++q)
          normal_pair(io.seed, io.sig, io.mem, (uint32_t)((i0 >> 1) + q), zz[2 * q], zz[2 * q + 1]);
      }
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const double xv = (i0 + k < n) ? io.x[i0 + k] : 0.0;
        v[k] = (T)(xv + zz[k] * io.sigma);
      }
      store8(dst + i0, v);
    } else if constexpr (MODE == PASS_FINAL) {
      T hv[RPL], rv[RPL];
      load8(src + i0, hv);
      load8(M.r + i0, rv);
#pragma unroll
      for (int k = 0; k < RPL; ++k) v[k] = rv[k] - hv[k];
      store8(M.r + i0, v);
      if (io.imf) {
        T* o = reinterpret_cast<T*>(io.imf);
#pragma unroll
        for (int k = 0; k < RPL; ++k) if (i0 + k < n) o[i0 + k] = hv[k];
      }
      if (io.acc) {
#pragma unroll
        for (int k = 0; k < RPL; ++k)
          if (i0 + k < n) io.acc[i0 + k] = (io.acc_first ? 0.0 : io.acc[i0 + k]) + (double)hv[k];
      }
    } else {
      // PASS_SIFT / PASS_LAST: mean envelope at the lane's samples first (only
      // the running sum is live during the evaluation), then h and r.
      T* const stage = S.stage + w * (RPL * WAVE);
#pragma unroll
      for (int ev = 0; ev < 2; ++ev) {
        const int byte = bc[ev][cb / 8 + ln];
        int tot;
        const int ex = wave_excl_small(__popc(byte), tot);
        const int base = S.pre[cur][ev][w] + crun[ev] + ex;
        crun[ev] += tot;
        envelope_chunk<T, NW>(S, M, cur, ev, w, base, byte, (int)i0, stage);
      }
      T msum[RPL];
#pragma unroll
      for (int k = 0; k < RPL; ++k) msum[k] = stage[k * WAVE + ln];
      T hv[RPL];
      load8(src + i0, hv);
      if constexpr (MODE == PASS_SIFT) {
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          v[k] = hv[k] - (T)0.5 * msum[k];
          if (io.want_sd && i0 + k < n) {
            const T dd = hv[k] - v[k];
            num += (double)(dd * dd);
            den += (double)(hv[k] * hv[k]);
          }
        }
        store8(M.h + i0, v);
      } else {
#pragma unroll
        for (int k = 0; k < RPL; ++k) hv[k] = hv[k] - (T)0.5 * msum[k];  // the IMF
        load8(M.r + i0, v);
#pragma unroll
        for (int k = 0; k < RPL; ++k) v[k] = v[k] - hv[k];                // new residual
        store8(M.r + i0, v);
        if (io.imf) {
          T* o = reinterpret_cast<T*>(io.imf);
#pragma unroll
          for (int k = 0; k < RPL; ++k) if (i0 + k < n) o[i0 + k] = hv[k];
        }
        if (io.acc) {
#pragma unroll
          for (int k = 0; k < RPL; ++k)
            if (i0 + k < n) io.acc[i0 + k] = (io.acc_first ? 0.0 : io.acc[i0 + k]) + (double)hv[k];
        }
      }
    }

    // ---- extrema of the new values (bit b <=> position b-1) ----
    T p1 = __shfl_up(v[RPL - 1], 1);
    T p2 = __shfl_up(v[RPL - 2], 1);
    if (ln == 0) {
      p1 = carry1;
      p2 = carry2;
    }
    carry1 = __shfl(v[RPL - 1], WAVE - 1);
    carry2 = __shfl(v[RPL - 2], WAVE - 1);
    const bool first = (cb == seg_lo) && (ln == 0);
    int bmax = 0, bmin = 0;
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int64_t b = i0 + k;
      const T vm1 = (k == 0) ? p2 : ((k == 1) ? p1 : v[k - 2]);
      const T v0 = (k == 0) ? p1 : v[k - 1];
      const T vp1 = v[k];
      const bool ok = (b >= 2) && (b <= n - 1) && !(first && k < 2);
      if (ok && v0 > vm1 && v0 > vp1) bmax |= 1 << k;
      if (ok && v0 < vm1 && v0 < vp1) bmin |= 1 << k;
    }
    bn[0][cb / 8 + ln] = (uint8_t)bmax;
    bn[1][cb / 8 + ln] = (uint8_t)bmin;
#pragma unroll
    for (int ev = 0; ev < 2; ++ev) {
      const int bits = ev == 0 ? bmax : bmin;
      int tot;
      const int ex = wave_excl_small(__popc(bits), tot);
      int o = run[ev] + ex;
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        if ((bits >> k) & 1) {
          const int q = (int)(i0 + k - 1);
          const T val = (k == 0) ? p1 : v[k - 1];
          if (o < kcap) {
            ktn[ev][SEG_FRONT + o] = q;
            kyn[ev][SEG_FRONT + o] = val;
          }
          if (io.trace) hs[ev] += mix64((unsigned long long)q);
          ++o;
        }
      }
      run[ev] += tot;
    }
  }

  if (io.trace) {
    const unsigned long long a = wave_sum_u64(hs[0]);
    const unsigned long long b = wave_sum_u64(hs[1]);
    if (ln == 0) {
      atomicAdd(&S.hsh[nxt][0], a);
      atomicAdd(&S.hsh[nxt][1], b);
    }
  }
  if (MODE == PASS_SIFT && io.want_sd) {
    num = wave_sum(num);
    den = wave_sum(den);
  }
  if (ln == 0) {
    S.cnt[nxt][0][w] = run[0];
    S.cnt[nxt][1][w] = run[1];
    S.beg[nxt][0][w] = SEG_FRONT;
    S.beg[nxt][1][w] = SEG_FRONT;
    S.red[0][w] = num;
    S.red[1][w] = den;
  }
}

// Positions seg_lo-1 and seg_lo of each segment need values of the previous
// segment; one thread per segment decides them after the pass and inserts
// them in front of the segment's knot list.
template <typename T, int NW>
__device__ __noinline__ void fixup(Shared<T, NW>& S, const Mem<T>& M, const T* vals, int set, int trace) {
  const int w = (int)threadIdx.x;
  if (w < 1 || w >= NW) return;
  const int64_t seg_lo = (int64_t)w * M.sl;
  if (seg_lo >= M.n) return;
  for (int ev = 0; ev < 2; ++ev) {
    for (int d = 0; d < 2; ++d) {
      const int64_t q = seg_lo - 1 + d;
      if (q < 1 || q > M.n - 2) continue;
      const T a = vals[q - 1], b = vals[q], c = vals[q + 1];
      const bool hit = ev == 0 ? (b > a && b > c) : (b < a && b < c);
      if (!hit) continue;
      const int bg = S.beg[set][ev][w] - 1;
      const int64_t base = (int64_t)(set * 2 + ev) * M.rstride + (int64_t)w * M.cap + bg;
      M.kt[base] = (int32_t)q;
      M.ky[base] = b;
      S.beg[set][ev][w] = bg;
      S.cnt[set][ev][w] += 1;
      uint8_t* bp = M.bits + (int64_t)(set * 2 + ev) * M.bstride + seg_lo / 8;
      *bp = (uint8_t)(*bp | (1 << d));
      if (trace) atomicAdd(&S.hsh[set][ev], mix64((unsigned long long)q));
    }
  }
}

template <typename T, int NW>
__device__ __noinline__ void prefix(Shared<T, NW>& S, int set) {
  for (int ev = 0; ev < 2; ++ev) {
    int a = 0;
    S.pre[set][ev][0] = 0;
    for (int w = 0; w < NW; ++w) {
      a += S.cnt[set][ev][w];
      S.pre[set][ev][w + 1] = a;
    }
  }
}

// Thread 0: envelope bookkeeping of set `cur` (knot counts, end extension,
// natural boundary M = 0 at the outermost knots).
template <typename T, int NW>
__device__ __noinline__ void setup_ends(Shared<T, NW>& S, const Mem<T>& M, int cur, int end_mode) {
  for (int ev = 0; ev < 2; ++ev) {
    const int p = S.pre[cur][ev][NW];
    S.p[ev] = p;
    S.nl[ev] = 0;
    S.nr[ev] = 0;
    S.P[ev] = p;
    S.cval[ev] = (T)0;
    int wh = 0;
    if (p == 1) {
      const Knot<T> k = fetch_knot(S, M, cur, ev, 0, wh);
      S.cval[ev] = k.y;
      continue;
    }
    if (p < 2) continue;
    if (end_mode == END_MIRROR && p >= 3) {
      // emd_ref.py:92-95: reflect the 2nd / penultimate extremum past each end
      const Knot<T> a = fetch_knot(S, M, cur, ev, 0, wh);
      const Knot<T> b = fetch_knot(S, M, cur, ev, 1, wh);
      int wh2 = NW - 1;
      const Knot<T> c = fetch_knot(S, M, cur, ev, p - 2, wh2);
      const Knot<T> d = fetch_knot(S, M, cur, ev, p - 1, wh2);
      S.vpos[ev][0] = a.t - (b.t - a.t);
      S.vy[ev][0] = b.y;
      S.vm[ev][0] = (T)0;
      S.vpos[ev][MAX_VIRT] = d.t + (d.t - c.t);
      S.vy[ev][MAX_VIRT] = c.y;
      S.vm[ev][MAX_VIRT] = (T)0;
      S.nl[ev] = 1;
      S.nr[ev] = 1;
      S.P[ev] = p + 2;
    } else {
      // natural: M = 0 at the first and last real knot
      const Knot<T> a = fetch_knot(S, M, cur, ev, 0, wh);
      int wh2 = NW - 1;
      const Knot<T> d = fetch_knot(S, M, cur, ev, p - 1, wh2);
      M.km[(int64_t)ev * M.rstride + a.code] = (T)0;
      M.km[(int64_t)ev * M.rstride + d.code] = (T)0;
    }
  }
}

template <typename T, int NW>
__device__ __forceinline__ void store_m(Shared<T, NW>& S, const Mem<T>& M, int env, int code,
                                        double val) {
  if (code >= 0) {
    M.km[(int64_t)env * M.rstride + code] = (T)val;
  } else {
    const int j = -1 - code;  // < MAX_VIRT: left, else right
    S.vm[env][j] = (T)val;
  }
}

// Decay-seamed block Thomas over both envelopes of set `cur`: owned blocks of
// B rows solved on the clipped window [o0-S, o1+S] with the couplings outside
// the window dropped (emd_gpu_v5_dsbt.py:51-74, 223-237); one window per
// thread.  B <= 0: one window per envelope (exact Thomas, emd_ref.py:50-65).
// Rows are built on the fly from the knots (emd_ref.py:31-48).
template <typename T, int NW>
__device__ __noinline__ void solve(Shared<T, NW>& S, const Mem<T>& M, int cur, int Bp, int Sp) {
  int nwin[2], nrow[2], Be[2];
  for (int ev = 0; ev < 2; ++ev) {
    const int p = S.p[ev];
    const int n = (p >= 2) ? S.P[ev] - 2 : 0;
    nrow[ev] = n;
    Be[ev] = (Bp <= 0 || Bp >= n) ? n : Bp;
    nwin[ev] = (n > 0) ? (n + Be[ev] - 1) / Be[ev] : 0;
  }
  const int total = nwin[0] + nwin[1];
  const int col = (int)threadIdx.x;
#pragma unroll 1
  for (int wi = (int)threadIdx.x; wi < total; wi += (int)blockDim.x) {
    const int ev = wi < nwin[0] ? 0 : 1;
    const int b = ev == 0 ? wi : wi - nwin[0];
    const int n = nrow[ev];
    const int o0 = b * Be[ev];
    int o1 = o0 + Be[ev];
    if (o1 > n) o1 = n;
    o1 -= 1;
    int w0 = o0 - Sp;
    if (w0 < 0) w0 = 0;
    int w1 = o1 + Sp;
    if (w1 > n - 1) w1 = n - 1;

    int wh = 0;
    Knot<double> ka, kb;
    {
      const Knot<T> a = fetch_knot(S, M, cur, ev, w0, wh);
      const Knot<T> c = fetch_knot(S, M, cur, ev, w0 + 1, wh);
      ka = {(double)a.t, (double)a.y, 0.0, a.code};
      kb = {(double)c.t, (double)c.y, 0.0, c.code};
    }
    double sl0 = (kb.y - ka.y) / (kb.t - ka.t);
    double cprev = 0.0, xprev = 0.0;
#pragma unroll 1
    for (int q = w0; q <= w1; ++q) {
      const Knot<T> cc = fetch_knot(S, M, cur, ev, q + 2, wh);
      const double ct = (double)cc.t, cy = (double)cc.y;
      const double H0 = kb.t - ka.t;
      const double H1 = ct - kb.t;
      const double d = 2.0 * (H0 + H1);
      const double l = (q >= 1) ? H0 : 0.0;
      const double u = (q <= n - 2) ? H1 : 0.0;
      const double sl1 = (cy - kb.y) / H1;
      const double r = 6.0 * (sl1 - sl0);
      const double den = d - l * cprev;
      const double c = (q == w1) ? 0.0 : u / den;
      const double xv = (r - l * xprev) / den;
      const int64_t row = (int64_t)(q - w0) * M.cols + col;
      M.wc[row] = c;
      M.wx[row] = xv;
      M.ws[row] = kb.code;
      cprev = c;
      xprev = xv;
      sl0 = sl1;
      ka = kb;
      kb = {ct, cy, 0.0, cc.code};
    }
    double xn = xprev;
    {
      const int64_t row = (int64_t)(w1 - w0) * M.cols + col;
      if (w1 >= o0 && w1 <= o1) store_m(S, M, ev, M.ws[row], xn);
    }
#pragma unroll 1
    for (int q = w1 - 1; q >= w0; --q) {
      const int64_t row = (int64_t)(q - w0) * M.cols + col;
      const double xi = M.wx[row] - M.wc[row] * xn;
      if (q >= o0 && q <= o1) store_m(S, M, ev, M.ws[row], xi);
      xn = xi;
    }
  }
}

// ------------------------------------------------------------------------
// the persistent kernel
// ------------------------------------------------------------------------
template <typename T, int NT>
__global__ __launch_bounds__(NT) void sift_kernel(SiftArgs A) {
  constexpr int NW = NT / WAVE;
  __shared__ Shared<T, NW> S;
  const int tid = (int)threadIdx.x;

  char* scr = A.scratch + (int64_t)blockIdx.x * A.scratch_stride;
  const ScratchLayout LY = make_layout((int)sizeof(T), A.npad, NW, A.cap, A.rows, A.cols);
  Mem<T> M;
  M.cap = A.cap;
  M.h = reinterpret_cast<T*>(scr + LY.h);
  M.r = reinterpret_cast<T*>(scr + LY.r);
  M.bits = reinterpret_cast<uint8_t*>(scr + LY.bits);
  M.bstride = A.npad / 8;
  M.kt = reinterpret_cast<int32_t*>(scr + LY.kt);
  M.ky = reinterpret_cast<T*>(scr + LY.ky);
  M.km = reinterpret_cast<T*>(scr + LY.km);
  M.rstride = (int64_t)NW * A.cap;
  M.wc = reinterpret_cast<double*>(scr + LY.wc);
  M.wx = reinterpret_cast<double*>(scr + LY.wx);
  M.ws = reinterpret_cast<int32_t*>(scr + LY.ws);
  M.cols = A.cols;

  const int trace = A.trace != nullptr;
  for (;;) {
    __syncthreads();
    if (tid == 0) S.task = atomicAdd(A.queue, 1);
    __syncthreads();
    const int tq = S.task;
    if (tq >= A.ntask) break;
    const int t = A.order[tq];
    const TaskDesc td = A.task[t];
    const SignalDesc sd = A.sig[td.signal];
    for (int m = tid; m < MAX_IMF_CAP; m += NT) S.acc_cnt[m] = 0;

    M.n = sd.n;
    M.sl = seg_len(sd.n, NW);

    for (int mem = td.m0; mem < td.m1; ++mem) {
      const int unit = sd.unit0 + mem;
      PassIO io = {};
      io.x = A.x + sd.x_off;
      io.z = (A.noise && sd.noise_off >= 0) ? A.noise + sd.noise_off + (int64_t)mem * sd.n : nullptr;
      io.sigma = sd.sigma;
      io.seed = A.seed;
      io.sig = (uint32_t)td.signal;
      io.mem = (uint32_t)mem;
      io.trace = trace;
      __syncthreads();
      if (tid == 0) {
        S.hsh[0][0] = S.hsh[0][1] = 0ull;
        S.hsh[1][0] = S.hsh[1][1] = 0ull;
      }
      __syncthreads();
      wave_pass<T, NW, PASS_INIT>(S, M, io, 1, 0);
      __syncthreads();
      fixup(S, M, M.r, 0, trace);
      __syncthreads();
      if (tid == 0) prefix(S, 0);
      __syncthreads();

      int cur = 0, m = 0, total = 0;
      for (;;) {
        const int cmax = S.pre[cur][0][NW], cmin = S.pre[cur][1][NW];
        if (cmax + cmin < A.min_extrema) break;
        if (m >= A.max_imf) break;
        int ns = 0;
        bool fused = false;
        for (int s = 0; s < A.max_sifts; ++s) {
          const int nmx = S.pre[cur][0][NW], nmn = S.pre[cur][1][NW];
          if (nmx < 1 || nmn < 1) break;
          __syncthreads();
          const int nxt = cur ^ 1;
          if (tid == 0) {
            setup_ends(S, M, cur, A.end_mode);
            S.hsh[nxt][0] = S.hsh[nxt][1] = 0ull;
            if (trace) {
              uint64_t* tr = A.trace + (((int64_t)unit * A.max_imf + m) * A.max_sifts + s) * 4;
              tr[0] = (uint64_t)nmx;
              tr[1] = (uint64_t)nmn;
              tr[2] = S.hsh[cur][0];
              tr[3] = S.hsh[cur][1];
            }
          }
          __syncthreads();
          solve(S, M, cur, A.B, A.S);
          __syncthreads();
          const bool last = (s == A.max_sifts - 1);
          io.src_is_r = (s == 0);
          if (last) {
            io.imf = A.member_imfs
                         ? (void*)(reinterpret_cast<T*>(A.member_imfs) + sd.mimf_off +
                                   ((int64_t)mem * A.max_imf + m) * sd.n)
                         : nullptr;
            io.acc = A.acc ? A.acc + td.acc_off + (int64_t)m * sd.n : nullptr;
            io.acc_first = (S.acc_cnt[m] == 0);
            wave_pass<T, NW, PASS_LAST>(S, M, io, cur, nxt);
          } else {
            io.want_sd = (A.tau > 0.0) && (s >= 1);
            wave_pass<T, NW, PASS_SIFT>(S, M, io, cur, nxt);
          }
          __syncthreads();
          fixup(S, M, last ? M.r : M.h, nxt, trace);
          __syncthreads();
          if (tid == 0) {
            prefix(S, nxt);
            double nu = 0.0, de = 0.0;
            for (int w = 0; w < NW; ++w) {
              nu += S.red[0][w];
              de += S.red[1][w];
            }
            S.sd = (de > 0.0) ? nu / de : 0.0;
          }
          __syncthreads();
          ++ns;
          cur = nxt;
          if (last) {
            fused = true;
            break;
          }
          if (A.tau > 0.0 && s >= 1 && S.sd < A.tau) break;
        }
        if (!fused) {
          const int nxt = cur ^ 1;
          __syncthreads();
          if (tid == 0) S.hsh[nxt][0] = S.hsh[nxt][1] = 0ull;
          __syncthreads();
          io.src_is_r = (ns == 0);
          io.imf = A.member_imfs
                       ? (void*)(reinterpret_cast<T*>(A.member_imfs) + sd.mimf_off +
                                 ((int64_t)mem * A.max_imf + m) * sd.n)
                       : nullptr;
          io.acc = A.acc ? A.acc + td.acc_off + (int64_t)m * sd.n : nullptr;
          io.acc_first = (S.acc_cnt[m] == 0);
          wave_pass<T, NW, PASS_FINAL>(S, M, io, cur, nxt);
          __syncthreads();
          fixup(S, M, M.r, nxt, trace);
          __syncthreads();
          if (tid == 0) prefix(S, nxt);
          __syncthreads();
          cur = nxt;
        }
        __syncthreads();
        if (tid == 0) {
          S.pm[m] = ns;
          S.acc_cnt[m] += 1;
        }
        total += ns;
        ++m;
        __syncthreads();
      }
      if (tid == 0) {
        A.nmode[unit] = m;
        A.nsift[unit] = total;
      }
      for (int j = tid; j < A.max_imf; j += NT) A.per_mode[(int64_t)unit * A.max_imf + j] = (j < m) ? S.pm[j] : 0;
    }
    __syncthreads();
    for (int j = tid; j < A.max_imf; j += NT) A.acc_cnt[(int64_t)t * A.max_imf + j] = S.acc_cnt[j];
  }
}

// Level-wise ensemble mean and residual (emd_ref.py:209-217): thread per
// sample; task accumulators of a signal are added in task (= member) order.
__global__ void mean_kernel(const SignalDesc* sig, int nsig, const int64_t* sig_start,
                            const TaskDesc* task, const double* acc, const int32_t* acc_cnt,
                            int max_imf, const double* x, double* avg, double* resid,
                            int32_t* nmodes, int64_t total_n) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= total_n) return;
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
    for (int t = sd.task0; t < sd.task0 + sd.ntask; ++t) {
      const int c = acc_cnt[(int64_t)t * max_imf + m];
      if (c > 0) {
        s = s + acc[task[t].acc_off + (int64_t)m * sd.n + i];
        cnt += c;
      }
    }
    if (cnt == 0) break;
    const double a = s / (double)cnt;
    avg[sd.avg_off + (int64_t)m * sd.n + i] = a;
    rs = (m == 0) ? a : rs + a;
    nm = m + 1;
  }
  resid[sd.x_off + i] = x[sd.x_off + i] - (nm > 0 ? rs : 0.0);
  if (i == 0) nmodes[lo] = nm;
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
template <typename T, int NT>
static hipError_t launch_sift_t(const SiftArgs& A, int grid, hipStream_t st) {
  hipLaunchKernelGGL((sift_kernel<T, NT>), dim3(grid), dim3(NT), 0, st, A);
  return hipGetLastError();
}

hipError_t launch_sift(const SiftArgs& A, int fp64, int nt, int grid, hipStream_t st) {
  if (fp64) {
    switch (nt) {
      case 256: return launch_sift_t<double, 256>(A, grid, st);
      case 512: return launch_sift_t<double, 512>(A, grid, st);
      case 1024: return launch_sift_t<double, 1024>(A, grid, st);
    }
  } else {
    switch (nt) {
      case 256: return launch_sift_t<float, 256>(A, grid, st);
      case 512: return launch_sift_t<float, 512>(A, grid, st);
      case 1024: return launch_sift_t<float, 1024>(A, grid, st);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t sift_occupancy(int fp64, int nt, int* blocks_per_cu) {
  const void* f = nullptr;
  if (fp64) {
    if (nt == 256) f = (const void*)sift_kernel<double, 256>;
    if (nt == 512) f = (const void*)sift_kernel<double, 512>;
    if (nt == 1024) f = (const void*)sift_kernel<double, 1024>;
  } else {
    if (nt == 256) f = (const void*)sift_kernel<float, 256>;
    if (nt == 512) f = (const void*)sift_kernel<float, 512>;
    if (nt == 1024) f = (const void*)sift_kernel<float, 1024>;
  }
  if (!f) return hipErrorInvalidValue;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, nt, 0);
}

hipError_t launch_mean(const SignalDesc* sig, int nsig, const int64_t* sig_start,
                       const TaskDesc* task, const double* acc, const int32_t* acc_cnt,
                       int max_imf, const double* x, double* avg, double* resid,
                       int32_t* nmodes, int64_t total_n, hipStream_t st) {
  if (total_n <= 0) return hipSuccess;
  const int bs = 256;
  const int64_t g = (total_n + bs - 1) / bs;
  hipLaunchKernelGGL(mean_kernel, dim3((unsigned)g), dim3(bs), 0, st, sig, nsig, sig_start,
                     task, acc, acc_cnt, max_imf, x, avg, resid, nmodes, total_n);
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
