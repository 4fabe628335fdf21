#pragma once
// Persistent EEMD sift engine for gfx950 (CDNA4, wave64): the kernel and its
// templated launchers.  Instantiated per precision by eemd_sift_f32.hip and
// eemd_sift_f64.hip (built with their own scheduling flags); the kernels
// shared by both precisions and the dispatchers are in eemd_sift.hip.
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
//   kty, km     knots per (set, envelope) by slot: (position, value) pairs and
//               spline second derivatives M in a separate array, so the
//               extrema pass writes pairs and the solve writes M densely.
//               Slots are grouped per segment region (the wave that owns a
//               signal segment writes its knots to its own region, so no
//               barrier is needed inside a pass); VSLOTS slots after the
//               regions hold the virtual end knots.
//   wcx         forward sweep {c', x'} of the window Thomas solves,
//               [rows][cols] pairs, column = the lane that owns the window
//
// FP64 mode follows the oracle's arithmetic exactly (no contraction: the file
// is compiled with -ffp-contract=off; IEEE divisions; the cube rounded once).
// FP32 mode stores samples and knots in float, evaluates envelopes in float
// with FMAs, and solves the spline systems in double.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "eemd_common.h"
#include "eemd_mlp.h"

namespace eemd {

// A knot value (real or virtual) and the slot of its record.  Virtual knots
// (end extension) live in the VSLOTS records after the segment regions.
template <typename T>
struct Knot {
  T t, y, m;
  int code;
};

template <typename T, int NW>
struct Shared {
  int32_t cnt[2][2][NW];     // knots per (set, envelope, segment)
  int32_t beg[2][2][NW];     // first used slot of each segment region
  int32_t pre[2][2][NW + 1]; // exclusive prefix of cnt over segments
  unsigned long long hsh[2][2];
  int32_t p[2], nl[2], nr[2], P[2];  // real, left / right virtual, total knots
  T cval[2];                         // constant envelope when p < 2
  T vt[2][2 * MAX_VIRT];             // positions of the virtual knots (interval search)
  int32_t acc_cnt[MAX_IMF_CAP];
  int32_t pm[MAX_IMF_CAP];
  int32_t task;
  double sd;
  // the SD stop's pairwise sums: leaves (offset, length) of the member's
  // partial last block in summation order, and a stack for combining them
  int32_t pw_off[PW_MAXPART], pw_len[PW_MAXPART];
  int32_t pw_n;
  int32_t pws_l[16], pws_n2[16], pws_st[16];
  double pws_a[16], pws_b[16];
  double pw_part[2];
  // per-wave staging of one envelope's new knots of a chunk (at most 4 per
  // lane: strict extrema are >= 2 samples apart), copied out coalesced
  typename std::conditional<std::is_same<T, float>::value, float2, double2>::type
      stage[NW][4 * WAVE + 1];  // + 1: write target of non-extremum positions
};

// Staging of the learned end extension (both ends of one sift).
struct LearnedLDS {
  double ls[2][2][MLP_CTX];  // [end][env][0]: distance of the nearest knot from the end
  __attribute__((aligned(16))) float x[2][MLP_IN];       // model input per end
  __attribute__((aligned(16))) float h[2][2][MLP_HID];   // hidden activations (ping-pong)
  float g[2], mu[2], sd[2];  // gap scale, value mean and spread per end
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

// Value of one (wave-uniform) lane broadcast to the wave with v_readlane on
// each 32-bit word, without a round trip through the LDS permute unit.
template <typename V>
__device__ __forceinline__ V lane_val(V v, int lane) {
  static_assert(sizeof(V) % 4 == 0, "32-bit words");
  union U {
    V v;
    int w[sizeof(V) / 4];
  } a, b;
  a.v = v;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(V) / 4); ++i) b.w[i] = __builtin_amdgcn_readlane(a.w[i], lane);
  return b.v;
}

// Value of the previous lane (lane 0 gets 0) by a whole-wave DPP shift
// (wave_shr:1), a VALU move instead of an LDS permute.
template <typename V>
__device__ __forceinline__ V lane_prev(V v) {
  static_assert(sizeof(V) % 4 == 0, "32-bit words");
  union U {
    V v;
    int w[sizeof(V) / 4];
  } a, b;
  a.v = v;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(V) / 4); ++i)
    b.w[i] = __builtin_amdgcn_update_dpp(0, a.w[i], 0x138, 0xf, 0xf, false);
  return b.v;
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
  uint8_t* bits;    // [2 sets][2 env][npad/8]
  int64_t bstride;  // npad/8
  // knots per (set, env), slot-indexed: positions and values as pairs,
  // spline second derivatives in their own array (dense writes by the
  // extrema pass and by the solve)
  typename std::conditional<std::is_same<T, float>::value, float2, double2>::type* kty;
  T* km;
  int64_t rstride;  // slots per (set, env)
  int64_t vbase;    // slot of virtual knot 0 (= NW*cap)
  void* wcx;        // window forward sweep {c', x'} [rows][cols] (T2)
  int64_t rows, cols;
  double* sdb;      // SD terms: (h - h')^2 [npad], h^2 [npad], leaf sums (tau > 0)
};

template <typename T>
using Pair = typename std::conditional<std::is_same<T, float>::value, float2, double2>::type;

// Index of slot `slot` of (set, env) in kty / km.
template <typename T>
__device__ __forceinline__ int64_t kix(const Mem<T>& M, int set, int env, int64_t slot) {
  return (int64_t)(set * 2 + env) * M.rstride + slot;
}

template <typename T>
__device__ __forceinline__ void put_knot(Pair<T>* p, int q, T val) {
  Pair<T> v;
  v.x = (T)q;
  v.y = val;
  *p = v;
}

// Slot of real knot g (0-based in time order) of (set, env); w is a segment
// hint updated in place.
template <typename T, int NW>
__device__ __forceinline__ int64_t real_slot(const Shared<T, NW>& S, const Mem<T>& M, int set,
                                             int env, int g, int& w) {
  const int32_t* pre = S.pre[set][env];
#pragma unroll 1
  while (w > 0 && g < pre[w]) --w;
#pragma unroll 1
  while (w < NW - 1 && g >= pre[w + 1]) ++w;
  return (int64_t)w * M.cap + S.beg[set][env][w] + (g - pre[w]);
}

// Slot of the knot with extended index e (virtual left knots first, then
// the real knots in time order, then virtual right knots).
template <typename T, int NW>
__device__ __forceinline__ int64_t ext_slot(const Shared<T, NW>& S, const Mem<T>& M, int set,
                                            int env, int e, int& w) {
  const int nl = S.nl[env];
  const int g = e - nl;
  if (g < 0) return M.vbase + e;
  const int p = S.p[env];
  if (g >= p) return M.vbase + MAX_VIRT + (g - p);
  return real_slot(S, M, set, env, g, w);
}

template <typename T, int NW>
__device__ __forceinline__ Knot<T> knot_at(const Shared<T, NW>& S, const Mem<T>& M, int set,
                                           int env, int e, int& w) {
  const int64_t slot = ext_slot(S, M, set, env, e, w);
  const int64_t ix = kix(M, set, env, slot);
  const Pair<T> v = M.kty[ix];
  Knot<T> k;
  k.t = v.x;
  k.y = v.y;
  k.m = M.km[ix];
  k.code = (int)slot;
  return k;
}

// Interval (extended index of its left knot) of sample i, given cl = number
// of real knots at positions < i: clip(count_ext(<i) - 1, 0, P-2), i.e.
// searchsorted-left minus one (emd_ref.py:70).
template <typename T, int NW>
__device__ __forceinline__ int ext_interval(const Shared<T, NW>& S, int env, int cl, int i, int nl,
                                            int nr, int p, int P) {
  // virtual knot positions ascend (left: outermost first; right: nearest
  // first) and usually lie outside the signal: the nearest one settles the
  // count, the others are compared only when it does not
  int ce = cl + nl;
  if (cl == 0 && nl > 0 && !(S.vt[env][nl - 1] < (T)i)) {
    ce = 0;
    for (int j = 0; j < nl - 1; ++j) ce += (S.vt[env][j] < (T)i) ? 1 : 0;
  }
  if (cl == p && nr > 0 && S.vt[env][MAX_VIRT] < (T)i) {
    ce += 1;
    for (int j = 1; j < nr; ++j) ce += (S.vt[env][MAX_VIRT + j] < (T)i) ? 1 : 0;
  }
  int e = ce - 1;
  e = e < 0 ? 0 : e;
  e = e > P - 2 ? P - 2 : e;
  return e;
}

// Cubic spline value on [k0, k1] at t (emd_ref.py:71-75), FP64 order of
// operations of the reference.
__device__ __forceinline__ double spline_val(double t0, double y0, double m0, double t1, double y1,
                                             double m1, double t) {
  const double hh = t1 - t0;
  const double a = (t1 - t) / hh;
  const double b = (t - t0) / hh;
  const double ca = cube_once(a), cb = cube_once(b);
  return (a * y0 + b * y1) + ((((ca - a) * m0) + ((cb - b) * m1)) * (hh * hh)) / 6.0;
}

// Candidate selection by a per-lane index with plain selects.  The
// candidates are separate scalars (never an array: the compiler would turn a
// select chain over an array into indexed scratch-memory loads).  j <= lim
// is known to the caller; a constant lim drops the impossible candidates.
template <int NC, typename V>
__device__ __forceinline__ V sel(int j, V a0, V a1, V a2, V a3, V a4, V a5, int lim = 5) {
  V r = a0;
  if constexpr (NC > 1) if (lim >= 1) r = (j == 1) ? a1 : r;
  if constexpr (NC > 2) if (lim >= 2) r = (j == 2) ? a2 : r;
  if constexpr (NC > 3) if (lim >= 3) r = (j == 3) ? a3 : r;
  if constexpr (NC > 4) if (lim >= 4) r = (j == 4) ? a4 : r;
  if constexpr (NC > 5) if (lim >= 5) r = (j == 5) ? a5 : r;
  return r;
}

// FP32: values of one envelope at samples i0..i0+7.  The lane loads its
// knots e0..e0+NC, turns each interval into Horner form (the expansion of
// emd_ref.py:71-75 in tau = t - t0: c1 = (y1-y0)/h - h(2 M0 + M1)/6,
// c2 = M0/2, c3 = (M1-M0)/(6h)) and selects the interval per sample.
struct Ivl4 {
  float t, y, c1, c2, c3;
};

__device__ __forceinline__ Ivl4 make_ivl(const Knot<float>& a, const Knot<float>& b) {
  const float hh = b.t - a.t;
  const float ih = __builtin_amdgcn_rcpf(hh);
  Ivl4 v;
  v.t = a.t;
  v.y = a.y;
  v.c1 = __builtin_fmaf(-hh * (1.0f / 6.0f), __builtin_fmaf(2.0f, a.m, b.m), (b.y - a.y) * ih);
  v.c2 = 0.5f * a.m;
  v.c3 = (b.m - a.m) * ih * (1.0f / 6.0f);
  return v;
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

// ------------------------------------------------------------------------
// software-pipelined sift pass
// ------------------------------------------------------------------------
// Per envelope and chunk: the lane's 8 extrema bits of the current set, the
// number of real knots before its first sample, the extended index of its
// first interval and up to six knots from there (loaded one chunk ahead).
template <typename T>
struct EnvChunk {
  int byte, base, e0;
  int fast;  // wave-uniform: every sample of the chunk lies between two real knots
  T t0, y0, m0, t1, y1, m1, t2, y2, m2, t3, y3, m3, t4, y4, m4, t5, y5, m5;
};

// The wave's own segment of one envelope of the current set (prefix, next
// prefix, first slot, knot count) and the envelope's real and left virtual
// knot counts.  Constant during a pass, so the chunk loop keeps them in
// registers instead of reading LDS.
struct SegCache {
  int pre, pre1, beg, cnt, p, nl;
};

template <typename T, int NW>
__device__ __forceinline__ SegCache seg_cache(const Shared<T, NW>& S, int set, int ev, int w) {
  SegCache c;
  c.pre = S.pre[set][ev][w];
  c.pre1 = (w < NW - 1) ? S.pre[set][ev][w + 1] : 0x7fffffff;
  c.beg = S.beg[set][ev][w];
  c.cnt = S.cnt[set][ev][w];
  c.p = S.p[ev];
  c.nl = S.nl[ev];
  return c;
}

// Stage 2 of the pipeline for one envelope: interval indices of the chunk
// and the loads of its knots.  Returns the wave-uniform candidate class
// (0: <= 2 intervals per lane, 1: <= 3, 2: <= 5).  Interior chunks (every
// sample between two real knots, decided for the whole wave from lanes 0 and
// 63) take the fast path: interval index = popcount of the bits, knot slots
// consecutive inside the segment run.
template <typename T, int NW>
__device__ __forceinline__ int env_prefetch(const Shared<T, NW>& S, const Mem<T>& M, int cur,
                                            int ev, int w, const SegCache& sc, int byte,
                                            int& crun, int i0, EnvChunk<T>& c) {
  int tot;
  const int ex = wave_excl_small(__popc(byte), tot);
  c.byte = byte;
  c.base = sc.pre + crun + ex;
  crun += tot;
  c.fast = 0;
  const int p = sc.p;
  if (p < 2) {
    c.e0 = 0;
    return 0;
  }
  const int nl = sc.nl;
  const int cl0 = c.base + (byte & 1);
  const int cl7 = c.base + __popc(byte);
  const int lo = __shfl(cl0, 0);
  const int hi = __shfl(cl7, WAVE - 1);
  // knot slots of one (set, envelope) region fit in 32 bits (signals have at
  // most 2^30 samples, a region about half as many slots): address the
  // region from its base pointers
  const int64_t rb = kix(M, cur, ev, 0);
  const Pair<T>* const kt = M.kty + rb;
  const T* const kmv = M.km + rb;
  auto ld = [&](int q, T& t, T& y, T& m) {
    const Pair<T> v = kt[(uint32_t)q];
    t = v.x;
    y = v.y;
    m = kmv[(uint32_t)q];
  };
  if (lo >= 1 && hi <= p - 1) {
    c.fast = 1;
    c.e0 = cl0 + nl - 1;
    const int span = cl7 - cl0;
    const int cls = (__ballot(span >= 3) != 0ull) ? 2 : ((__ballot(span >= 2) != 0ull) ? 1 : 0);
    const int g0 = cl0 - 1;
    int wg = w;
    int s0;
    int rem;
    if (g0 >= sc.pre && g0 < sc.pre1) {
      s0 = w * (int)M.cap + sc.beg + (g0 - sc.pre);
      rem = sc.cnt - (g0 - sc.pre);
    } else {
      s0 = (int)real_slot(S, M, cur, ev, g0, wg);
      rem = S.cnt[cur][ev][wg] - (g0 - S.pre[cur][ev][wg]);
    }
    const int need = span + 2;  // knots e0 .. e0+span+1, all real
    if (rem >= need) {
      ld(s0, c.t0, c.y0, c.m0);
      ld(s0 + 1, c.t1, c.y1, c.m1);
      if (need > 2) ld(s0 + 2, c.t2, c.y2, c.m2);
      if (cls >= 1 && need > 3) ld(s0 + 3, c.t3, c.y3, c.m3);
      if (cls >= 2 && need > 4) ld(s0 + 4, c.t4, c.y4, c.m4);
      if (cls >= 2 && need > 5) ld(s0 + 5, c.t5, c.y5, c.m5);
    } else {
      int wh = wg;
      auto at = [&](int k) { return (int)real_slot(S, M, cur, ev, g0 + k, wh); };
      ld(at(0), c.t0, c.y0, c.m0);
      ld(at(1), c.t1, c.y1, c.m1);
      if (need > 2) ld(at(2), c.t2, c.y2, c.m2);
      if (cls >= 1 && need > 3) ld(at(3), c.t3, c.y3, c.m3);
      if (cls >= 2 && need > 4) ld(at(4), c.t4, c.y4, c.m4);
      if (cls >= 2 && need > 5) ld(at(5), c.t5, c.y5, c.m5);
    }
    return cls;
  }
  const int nr = S.nr[ev], P = S.P[ev];
  c.e0 = ext_interval(S, ev, cl0, i0, nl, nr, p, P);
  const int e7 = ext_interval(S, ev, cl7, i0 + RPL - 1, nl, nr, p, P);
  const int span = e7 - c.e0;
  const int cls = (__ballot(span >= 3) != 0ull) ? 2 : ((__ballot(span >= 2) != 0ull) ? 1 : 0);
  int wh = w;
  const int last = P - 1;
  auto at = [&](int k) {
    int e = c.e0 + k;
    e = e > last ? last : e;
    return (int)ext_slot(S, M, cur, ev, e, wh);
  };
  ld(at(0), c.t0, c.y0, c.m0);
  ld(at(1), c.t1, c.y1, c.m1);
  ld(at(2), c.t2, c.y2, c.m2);
  if (cls >= 1) ld(at(3), c.t3, c.y3, c.m3);
  if (cls >= 2) {
    ld(at(4), c.t4, c.y4, c.m4);
    ld(at(5), c.t5, c.y5, c.m5);
  }
  return cls;
}

// Stage 2, fixed-size form (FP64): as env_prefetch, but every path issues
// the same six knot loads (surplus ones repeat a valid slot).  With a fixed
// number of memory operations per chunk, the waits for one chunk's data do
// not extend to the loads already in flight for the next.  For FP32 the
// surplus loads cost more than the waits they save.  Returns the
// wave-uniform candidate class (0: <= 2 intervals per lane, 1: <= 3, 2: <= 5).
// Interior chunks (every sample between two real knots, decided for the
// whole wave from lanes 0 and 63) take the fast path: interval index =
// popcount of the bits, knot slots consecutive inside the segment run.
template <typename T, int NW>
__device__ __forceinline__ int env_prefetch_fixed(const Shared<T, NW>& S, const Mem<T>& M, int cur,
                                                  int ev, int w, const SegCache& sc, int byte,
                                                  int& crun, int i0, EnvChunk<T>& c) {
  int tot;
  const int ex = wave_excl_small(__popc(byte), tot);
  c.byte = byte;
  c.base = sc.pre + crun + ex;
  crun += tot;
  c.fast = 0;
  c.e0 = 0;
  const int p = sc.p;
  // knot slots of one (set, envelope) region fit in 32 bits (signals have at
  // most 2^30 samples, a region about half as many slots)
  const int64_t rb = kix(M, cur, ev, 0);
  const Pair<T>* const kt = M.kty + rb;
  const T* const kmv = M.km + rb;
  int x0 = 0, x1 = 0, x2 = 0, x3 = 0, x4 = 0, x5 = 0;  // knot slots
  int cls = 0;
  if (p >= 2) {
    const int nl = sc.nl;
    const int cl0 = c.base + (byte & 1);
    const int cl7 = c.base + __popc(byte);
    const int lo = lane_val(cl0, 0);
    const int hi = lane_val(cl7, WAVE - 1);
    if (lo >= 1 && hi <= p - 1) {
      c.fast = 1;
      c.e0 = cl0 + nl - 1;
      const int span = cl7 - cl0;
      cls = (__ballot(span >= 3) != 0ull) ? 2 : ((__ballot(span >= 2) != 0ull) ? 1 : 0);
      const int g0 = cl0 - 1;
      int wg = w;
      int s0;
      int rem;
      if (g0 >= sc.pre && g0 < sc.pre1) {
        s0 = w * (int)M.cap + sc.beg + (g0 - sc.pre);
        rem = sc.cnt - (g0 - sc.pre);
      } else {
        s0 = (int)real_slot(S, M, cur, ev, g0, wg);
        rem = S.cnt[cur][ev][wg] - (g0 - S.pre[cur][ev][wg]);
      }
      const int lastk = span + 1;  // knots g0 .. g0+span+1 are all real
      auto kk = [&](int k) { return k < lastk ? k : lastk; };
      if (rem > lastk) {
        x0 = s0;
        x1 = s0 + kk(1);
        x2 = s0 + kk(2);
        x3 = s0 + kk(3);
        x4 = s0 + kk(4);
        x5 = s0 + kk(5);
      } else {
        int wh = wg;
        x0 = s0;
        x1 = (int)real_slot(S, M, cur, ev, g0 + kk(1), wh);
        x2 = (int)real_slot(S, M, cur, ev, g0 + kk(2), wh);
        x3 = (int)real_slot(S, M, cur, ev, g0 + kk(3), wh);
        x4 = (int)real_slot(S, M, cur, ev, g0 + kk(4), wh);
        x5 = (int)real_slot(S, M, cur, ev, g0 + kk(5), wh);
      }
    } else {
      const int nr = S.nr[ev], P = S.P[ev];
      c.e0 = ext_interval(S, ev, cl0, i0, nl, nr, p, P);
      const int e7 = ext_interval(S, ev, cl7, i0 + RPL - 1, nl, nr, p, P);
      const int span = e7 - c.e0;
      cls = (__ballot(span >= 3) != 0ull) ? 2 : ((__ballot(span >= 2) != 0ull) ? 1 : 0);
      int wh = w;
      const int last = P - 1;
      auto at = [&](int k) {
        int e = c.e0 + k;
        e = e > last ? last : e;
        return (int)ext_slot(S, M, cur, ev, e, wh);
      };
      x0 = at(0);
      x1 = at(1);
      x2 = at(2);
      x3 = at(3);
      x4 = at(4);
      x5 = at(5);
    }
  }
  auto ld = [&](int q, T& t, T& y, T& m) {
    const Pair<T> v = kt[(uint32_t)q];
    t = v.x;
    y = v.y;
    m = kmv[(uint32_t)q];
  };
  ld(x0, c.t0, c.y0, c.m0);
  ld(x1, c.t1, c.y1, c.m1);
  ld(x2, c.t2, c.y2, c.m2);
  ld(x3, c.t3, c.y3, c.m3);
  ld(x4, c.t4, c.y4, c.m4);
  ld(x5, c.t5, c.y5, c.m5);
  return cls;
}

// Stage 3: values of the envelope at samples i0..i0+7 (written for ev 0,
// added for ev 1).
template <int NC, bool FAST, typename T, int NW>
__device__ __forceinline__ void env_eval(const Shared<T, NW>& S, int ev, const EnvChunk<T>& c,
                                         int i0, T (&acc)[RPL]) {
  const int nl = S.nl[ev], nr = S.nr[ev], p = S.p[ev], P = S.P[ev];
  if constexpr (std::is_same<T, double>::value) {
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int j = FAST ? __popc(c.byte & ((2u << k) - 2u))
                         : ext_interval(S, ev, c.base + __popc(c.byte & ((2u << k) - 1u)), i0 + k,
                                        nl, nr, p, P) - c.e0;
      const int j1 = j + 1;
      // FAST: j counts the knots at positions 1..k, i.e. the strict extrema
      // of one envelope among k consecutive samples, so j <= (k + 1) / 2
      const int lj = FAST ? (k + 1) / 2 : 5, lj1 = FAST ? (k + 1) / 2 + 1 : 5;
      const double v = spline_val(sel<NC + 1>(j, c.t0, c.t1, c.t2, c.t3, c.t4, c.t5, lj),
                                  sel<NC + 1>(j, c.y0, c.y1, c.y2, c.y3, c.y4, c.y5, lj),
                                  sel<NC + 1>(j, c.m0, c.m1, c.m2, c.m3, c.m4, c.m5, lj),
                                  sel<NC + 1>(j1, c.t0, c.t1, c.t2, c.t3, c.t4, c.t5, lj1),
                                  sel<NC + 1>(j1, c.y0, c.y1, c.y2, c.y3, c.y4, c.y5, lj1),
                                  sel<NC + 1>(j1, c.m0, c.m1, c.m2, c.m3, c.m4, c.m5, lj1),
                                  (double)(i0 + k));
      acc[k] = ev == 0 ? v : acc[k] + v;
    }
  } else {
    const Knot<float> k0{c.t0, c.y0, c.m0, 0}, k1{c.t1, c.y1, c.m1, 0}, k2{c.t2, c.y2, c.m2, 0};
    const Knot<float> k3{c.t3, c.y3, c.m3, 0}, k4{c.t4, c.y4, c.m4, 0}, k5{c.t5, c.y5, c.m5, 0};
    const Ivl4 v0 = make_ivl(k0, k1);
    const Ivl4 v1 = make_ivl(k1, k2);
    const Ivl4 v2 = NC > 2 ? make_ivl(k2, k3) : v1;
    const Ivl4 v3 = NC > 3 ? make_ivl(k3, k4) : v1;
    const Ivl4 v4 = NC > 4 ? make_ivl(k4, k5) : v1;
    if constexpr (FAST) {
      // interior chunk: every knot sits on a sample, so t_j - i0 and each
      // tau = (i0 - t_j) + k are exact; evaluate every candidate interval
      // and select the value (one select per candidate instead of five)
      const float fi0 = (float)i0;
      const float b0 = fi0 - v0.t, b1 = fi0 - v1.t, b2 = fi0 - v2.t, b3 = fi0 - v3.t,
                  b4 = fi0 - v4.t;
      auto horner = [](float tau, const Ivl4& iv) {
        return __builtin_fmaf(tau, __builtin_fmaf(tau, __builtin_fmaf(tau, iv.c3, iv.c2), iv.c1),
                              iv.y);
      };
      // j counts the knots at positions 1..k, the strict extrema of one
      // envelope among k consecutive samples (never two adjacent), so j <=
      // (k + 1) / 2 (k is a constant after unrolling)
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int j = __popc(c.byte & ((2u << k) - 2u));
        const float fk = (float)k;
        float v = horner(b0 + fk, v0);
        const int jmax = (k + 1) / 2;
        if (jmax >= 1) v = (j == 1) ? horner(b1 + fk, v1) : v;
        if constexpr (NC > 2) if (jmax >= 2) v = (j == 2) ? horner(b2 + fk, v2) : v;
        if constexpr (NC > 3) if (jmax >= 3) v = (j == 3) ? horner(b3 + fk, v3) : v;
        if constexpr (NC > 4) if (jmax >= 4) v = (j == 4) ? horner(b4 + fk, v4) : v;
        acc[k] = ev == 0 ? v : acc[k] + v;
      }
      return;
    }
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int j = FAST ? __popc(c.byte & ((2u << k) - 2u))
                         : ext_interval(S, ev, c.base + __popc(c.byte & ((2u << k) - 1u)), i0 + k,
                                        nl, nr, p, P) - c.e0;
      const float t = sel<NC>(j, v0.t, v1.t, v2.t, v3.t, v4.t, v4.t);
      const float y = sel<NC>(j, v0.y, v1.y, v2.y, v3.y, v4.y, v4.y);
      const float c1 = sel<NC>(j, v0.c1, v1.c1, v2.c1, v3.c1, v4.c1, v4.c1);
      const float c2 = sel<NC>(j, v0.c2, v1.c2, v2.c2, v3.c2, v4.c2, v4.c2);
      const float c3 = sel<NC>(j, v0.c3, v1.c3, v2.c3, v3.c3, v4.c3, v4.c3);
      const float tau = (float)(i0 + k) - t;
      const float v = __builtin_fmaf(tau, __builtin_fmaf(tau, __builtin_fmaf(tau, c3, c2), c1), y);
      acc[k] = ev == 0 ? v : acc[k] + v;
    }
  }
}

template <typename T, int NW>
__device__ __forceinline__ void env_eval_any(const Shared<T, NW>& S, int ev, const SegCache& sc,
                                             int cls, const EnvChunk<T>& c, int i0,
                                             T (&acc)[RPL]) {
  if (sc.p < 2) {
#pragma unroll
    for (int k = 0; k < RPL; ++k) acc[k] = ev == 0 ? S.cval[ev] : acc[k] + S.cval[ev];
    return;
  }
  if (c.fast) {
    if (cls == 2) env_eval<5, true>(S, ev, c, i0, acc);
    else if (cls == 1) env_eval<3, true>(S, ev, c, i0, acc);
    else env_eval<2, true>(S, ev, c, i0, acc);
  } else {
    if (cls == 2) env_eval<5, false>(S, ev, c, i0, acc);
    else if (cls == 1) env_eval<3, false>(S, ev, c, i0, acc);
    else env_eval<2, false>(S, ev, c, i0, acc);
  }
}

// Extrema of the new values v (bit b <=> position b-1), written to the
// bitmaps of set nxt and compacted, in time order, into the wave's segment
// region (emd_ref.py:84-87).  p1/p2 come from the lane before (or the carry
// of the previous chunk); the first two positions of a segment are decided
// by fixup() after the pass.
template <typename T, int NW>
__device__ __forceinline__ void detect_compact(Shared<T, NW>& S, const Mem<T>& M, int nxt,
                                               int64_t cb, int64_t seg_lo,
                                               const T (&v)[RPL], T& carry1, T& carry2,
                                               int (&run)[2], unsigned long long (&hs)[2],
                                               int trace, int64_t seg_slot) {
  const int ln = lane_id();
  // sample positions fit in 32 bits (signals have at most 2^30 samples)
  const int n = (int)M.n;
  const int i0 = (int)cb + ln * RPL;
  const int kcap = (int)(M.cap - SEG_FRONT);
  uint8_t* const bn0 = M.bits + (int64_t)(nxt * 2 + 0) * M.bstride;
  uint8_t* const bn1 = M.bits + (int64_t)(nxt * 2 + 1) * M.bstride;
  T p1 = lane_prev(v[RPL - 1]);
  T p2 = lane_prev(v[RPL - 2]);
  if (ln == 0) {
    p1 = carry1;
    p2 = carry2;
  }
  // FP64: v_readlane broadcast (measured faster there); FP32: the permute
  if constexpr (std::is_same<T, double>::value) {
    carry1 = lane_val(v[RPL - 1], WAVE - 1);
    carry2 = lane_val(v[RPL - 2], WAVE - 1);
  } else {
    carry1 = __shfl(v[RPL - 1], WAVE - 1);
    carry2 = __shfl(v[RPL - 2], WAVE - 1);
  }
  const bool first = (cb == seg_lo) && (ln == 0);
  int bmax = 0, bmin = 0;
#pragma unroll
  for (int k = 0; k < RPL; ++k) {
    const int b = i0 + k;
    const T vm1 = (k == 0) ? p2 : ((k == 1) ? p1 : v[k - 2]);
    const T v0 = (k == 0) ? p1 : v[k - 1];
    const T vp1 = v[k];
    const bool ok = (b >= 2) && (b <= n - 1) && !(first && k < 2);
    if (ok && v0 > vm1 && v0 > vp1) bmax |= 1 << k;
    if (ok && v0 < vm1 && v0 < vp1) bmin |= 1 << k;
  }
  bn0[cb / 8 + ln] = (uint8_t)bmax;
  bn1[cb / 8 + ln] = (uint8_t)bmin;
  Pair<T>* const st = S.stage[threadIdx.x / WAVE];
#pragma unroll
  for (int ev = 0; ev < 2; ++ev) {
    const int bits = ev == 0 ? bmax : bmin;
    int tot;
    const int ex = wave_excl_small(__popc(bits), tot);
    if (tot == 0) continue;  // wave-uniform
    // branch-free: every position writes, non-extrema to the spare entry
    int o = ex;
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int bit = (bits >> k) & 1;
      Pair<T> kv;
      kv.x = (T)(i0 + k - 1);
      kv.y = (k == 0) ? p1 : v[k - 1];
      st[bit ? o : 4 * WAVE] = kv;
      o += bit;
    }
    if (trace) {
#pragma unroll
      for (int k = 0; k < RPL; ++k)
        if ((bits >> k) & 1) hs[ev] += mix64((unsigned long long)(i0 + k - 1));
    }
    // the wave's own LDS writes above are complete before its reads below
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    Pair<T>* const dst = M.kty + kix(M, nxt, ev, seg_slot + run[ev]);
    for (int k = ln; k < tot; k += WAVE)
      if (run[ev] + k < kcap) dst[k] = st[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    run[ev] += tot;
  }
}

template <typename T, int NW>
__device__ __forceinline__ void pass_epilogue(Shared<T, NW>& S, int nxt, int w,
                                              const int (&run)[2],
                                              const unsigned long long (&hs)[2], int trace) {
  const int ln = lane_id();
  if (trace) {
    const unsigned long long a = wave_sum_u64(hs[0]);
    const unsigned long long b = wave_sum_u64(hs[1]);
    if (ln == 0) {
      atomicAdd(&S.hsh[nxt][0], a);
      atomicAdd(&S.hsh[nxt][1], b);
    }
  }
  if (ln == 0) {
    S.cnt[nxt][0][w] = run[0];
    S.cnt[nxt][1][w] = run[1];
    S.beg[nxt][0][w] = SEG_FRONT;
    S.beg[nxt][1][w] = SEG_FRONT;
  }
}

// SIFT / LAST pass, software-pipelined: while chunk c is evaluated, the
// bitmap bytes of chunk c+2 and the knots, h (and r) of chunk c+1 are in
// flight.  FP64 issues the group of loads on every iteration (past the end
// of the segment it rereads the current chunk and the result goes unused),
// so its size is the same on every path and waiting for chunk c's data
// never waits for chunk c+1's; FP32 skips the group past the end.
template <typename T, int NW, bool LAST>
__device__ __forceinline__ void wave_sift_pipe(Shared<T, NW>& S, const Mem<T>& M, const PassIO& io,
                                               int cur, int nxt) {
  const int ln = lane_id();
  const int w = (int)(threadIdx.x / WAVE);
  const int64_t n = M.n;
  const int64_t seg_lo = (int64_t)w * M.sl;
  int64_t seg_hi = seg_lo + M.sl;
  if (seg_hi > n) seg_hi = n;
  int run[2] = {0, 0};
  unsigned long long hs[2] = {0ull, 0ull};
  T carry1 = (T)0, carry2 = (T)0;
  constexpr bool FIXED = std::is_same<T, double>::value;
  const bool want_sd = !LAST && io.want_sd;
  const T* const src = io.src_is_r ? M.r : M.h;
  const uint8_t* const bc0 = M.bits + (int64_t)(cur * 2 + 0) * M.bstride;
  const uint8_t* const bc1 = M.bits + (int64_t)(cur * 2 + 1) * M.bstride;
  const int64_t seg_slot = (int64_t)w * M.cap + SEG_FRONT;

  if (seg_lo < seg_hi) {
    int crun0 = 0, crun1 = 0;
    int64_t cb = seg_lo;
    // sample positions fit in 32 bits (signals have at most 2^30 samples)
    const int i00 = (int)cb + ln * RPL;
    const SegCache sc0 = seg_cache(S, cur, 0, w), sc1 = seg_cache(S, cur, 1, w);
    int fb0 = 0, fb1 = 0;
    if (FIXED || cb + CHUNK < seg_hi) {
      const int64_t cb1 = (cb + CHUNK < seg_hi) ? cb + CHUNK : cb;
      fb0 = bc0[cb1 / 8 + ln];
      fb1 = bc1[cb1 / 8 + ln];
    }
    EnvChunk<T> A0, A1;
    int cls0, cls1;
    if constexpr (FIXED) {
      cls0 = env_prefetch_fixed(S, M, cur, 0, w, sc0, (int)bc0[cb / 8 + ln], crun0, i00, A0);
      cls1 = env_prefetch_fixed(S, M, cur, 1, w, sc1, (int)bc1[cb / 8 + ln], crun1, i00, A1);
    } else {
      cls0 = env_prefetch(S, M, cur, 0, w, sc0, (int)bc0[cb / 8 + ln], crun0, i00, A0);
      cls1 = env_prefetch(S, M, cur, 1, w, sc1, (int)bc1[cb / 8 + ln], crun1, i00, A1);
    }
    T hv[RPL], rv[RPL];
    load8(src + (uint32_t)i00, hv);
    if constexpr (LAST) load8(M.r + (uint32_t)i00, rv);
#pragma unroll 1
    for (; cb < seg_hi; cb += CHUNK) {
      const int i0 = (int)cb + ln * RPL;
      const int ni = (int)n;
      // FP64 LAST: the accumulator values of chunk c are read before the
      // group below, so that waiting for them does not wait for the group
      double av[RPL];
      if constexpr (FIXED && LAST) {
        const bool rd = io.acc && !io.acc_first;
#pragma unroll
        for (int k = 0; k < RPL; ++k) av[k] = (rd && i0 + k < ni) ? io.acc[(uint32_t)(i0 + k)] : 0.0;
      }
      EnvChunk<T> B0, B1;
      int clsB0 = 0, clsB1 = 0;
      int gb0 = 0, gb1 = 0;
      T hn[RPL], rn[RPL];
      if constexpr (FIXED) {
        // chunk c+1 (c+2 for the bitmaps), or chunk c again past the end
        const int64_t d1 = (cb + CHUNK < seg_hi) ? CHUNK : 0;
        const int64_t d2 = (cb + 2 * CHUNK < seg_hi) ? 2 * CHUNK : d1;
        gb0 = bc0[(cb + d2) / 8 + ln];
        gb1 = bc1[(cb + d2) / 8 + ln];
        clsB0 = env_prefetch_fixed(S, M, cur, 0, w, sc0, fb0, crun0, i0 + (int)d1, B0);
        clsB1 = env_prefetch_fixed(S, M, cur, 1, w, sc1, fb1, crun1, i0 + (int)d1, B1);
        load8(src + (uint32_t)(i0 + (int)d1), hn);
        if constexpr (LAST) load8(M.r + (uint32_t)(i0 + (int)d1), rn);
      } else {
        const bool has1 = cb + CHUNK < seg_hi;
        const bool has2 = cb + 2 * CHUNK < seg_hi;
        if (has2) {
          gb0 = bc0[(cb + 2 * CHUNK) / 8 + ln];
          gb1 = bc1[(cb + 2 * CHUNK) / 8 + ln];
        }
        if (has1) {
          clsB0 = env_prefetch(S, M, cur, 0, w, sc0, fb0, crun0, i0 + CHUNK, B0);
          clsB1 = env_prefetch(S, M, cur, 1, w, sc1, fb1, crun1, i0 + CHUNK, B1);
          load8(src + (uint32_t)(i0 + CHUNK), hn);
          if constexpr (LAST) load8(M.r + (uint32_t)(i0 + CHUNK), rn);
        }
      }
      T msum[RPL];
      env_eval_any(S, 0, sc0, cls0, A0, i0, msum);
      env_eval_any(S, 1, sc1, cls1, A1, i0, msum);
      T v[RPL];
      if constexpr (!LAST) {
#pragma unroll
        for (int k = 0; k < RPL; ++k) v[k] = hv[k] - (T)0.5 * msum[k];
        if (want_sd) {
          // the terms of the SD sums, summed afterwards in numpy's order
          const int64_t np = (int64_t)M.bstride * 8;
#pragma unroll
          for (int k = 0; k < RPL; ++k)
            if (i0 + k < ni) {
              const T dd = hv[k] - v[k];
              M.sdb[i0 + k] = (double)(dd * dd);
              M.sdb[np + i0 + k] = (double)(hv[k] * hv[k]);
            }
        }
        store8(M.h + (uint32_t)i0, v);
      } else {
        T imf[RPL];
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          imf[k] = hv[k] - (T)0.5 * msum[k];
          v[k] = rv[k] - imf[k];
        }
        store8(M.r + (uint32_t)i0, v);
        if (io.imf) {
          T* o = reinterpret_cast<T*>(io.imf);
#pragma unroll
          for (int k = 0; k < RPL; ++k) if (i0 + k < ni) o[(uint32_t)(i0 + k)] = imf[k];
        }
        if (io.acc) {
#pragma unroll
          for (int k = 0; k < RPL; ++k)
            if (i0 + k < ni) {
              const uint32_t q = (uint32_t)(i0 + k);
              if constexpr (FIXED) io.acc[q] = av[k] + (double)imf[k];
              else io.acc[q] = (io.acc_first ? 0.0 : io.acc[q]) + (double)imf[k];
            }
        }
      }
      detect_compact<T, NW>(S, M, nxt, cb, seg_lo, v, carry1, carry2, run, hs, io.trace, seg_slot);
      A0 = B0;
      A1 = B1;
      cls0 = clsB0;
      cls1 = clsB1;
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        hv[k] = hn[k];
        if constexpr (LAST) rv[k] = rn[k];
      }
      fb0 = gb0;
      fb1 = gb1;
    }
  }
  pass_epilogue(S, nxt, w, run, hs, io.trace);
}

// INIT / FINAL pass: one wave streams its segment, produces the new residual
// (noise added, or the IMF subtracted), writes it and compacts its extrema
// into set `nxt`.
template <typename T, int NW, int MODE>
__device__ __forceinline__ void wave_pass(Shared<T, NW>& S, const Mem<T>& M, const PassIO& io, int cur, int nxt) {
  static_assert(MODE == PASS_INIT || MODE == PASS_FINAL, "sift passes use wave_sift_pipe");
  const int ln = lane_id();
  const int w = (int)(threadIdx.x / WAVE);
  const int64_t n = M.n;
  const int64_t seg_lo = (int64_t)w * M.sl;
  int64_t seg_hi = seg_lo + M.sl;
  if (seg_hi > n) seg_hi = n;
  int run[2] = {0, 0};
  unsigned long long hs[2] = {0ull, 0ull};
  T carry1 = (T)0, carry2 = (T)0;
  const T* const src = io.src_is_r ? M.r : M.h;

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
      store8(M.r + i0, v);
    } else {
      T hv[RPL];
      load8(src + i0, hv);
      load8(M.r + i0, v);
#pragma unroll
      for (int k = 0; k < RPL; ++k) v[k] = v[k] - hv[k];
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
    detect_compact<T, NW>(S, M, nxt, cb, seg_lo, v, carry1, carry2, run, hs, io.trace,
                          (int64_t)w * M.cap + SEG_FRONT);
  }
  pass_epilogue(S, nxt, w, run, hs, io.trace);
}

// ------------------------------------------------------------------------
// SD stop test in numpy's summation order (emd_ref.py:141-142 sums with
// np.sum): the pass stores the per-sample terms, these routines add them
// the way numpy's pairwise summation does.
// ------------------------------------------------------------------------
// Sum of one leaf of n <= PW_LEAF elements: in order when n < 8; otherwise
// eight running sums of stride-8 elements, combined as
// ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then the remaining elements in order.
__device__ __forceinline__ double pw_leaf(const double* a, int n) {
  if (n < 8) {
    double r = 0.0;
    for (int i = 0; i < n; ++i) r += a[i];
    return r;
  }
  double r0 = a[0], r1 = a[1], r2 = a[2], r3 = a[3], r4 = a[4], r5 = a[5], r6 = a[6], r7 = a[7];
  int i = 8;
  for (; i < n - (n % 8); i += 8) {
    r0 += a[i];
    r1 += a[i + 1];
    r2 += a[i + 2];
    r3 += a[i + 3];
    r4 += a[i + 4];
    r5 += a[i + 5];
    r6 += a[i + 6];
    r7 += a[i + 7];
  }
  double res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
  for (; i < n; ++i) res += a[i];
  return res;
}

// Split point of a range of l > PW_LEAF elements: half, rounded down to a
// multiple of 8.
__device__ __forceinline__ int pw_split(int l) {
  const int n2 = l / 2;
  return n2 - n2 % 8;
}

// Thread 0, once per member: the leaves of the partial last block (n mod
// PW_BLOCK elements) in summation order (depth-first, left half first).
template <typename T, int NW>
__device__ void pw_enumerate(Shared<T, NW>& S, int64_t n) {
  const int m = (int)(n % PW_BLOCK);
  int cnt = 0;
  if (m > 0) {
    int sp = 0;
    S.pws_l[0] = m;
    S.pws_n2[0] = 0;  // offset of the range
    sp = 1;
    while (sp > 0) {
      --sp;
      const int l = S.pws_l[sp], o = S.pws_n2[sp];
      if (l <= PW_LEAF) {
        S.pw_off[cnt] = o;
        S.pw_len[cnt] = l;
        ++cnt;
        continue;
      }
      const int n2 = pw_split(l);
      S.pws_l[sp] = l - n2;  // right half, taken after the left one
      S.pws_n2[sp] = o + n2;
      ++sp;
      S.pws_l[sp] = n2;
      S.pws_n2[sp] = o;
      ++sp;
    }
  }
  S.pw_n = cnt;
}

// All threads, after a pass that stored the SD terms: S.sd = num / den (0
// when den is 0) with num and den summed in numpy's order.
template <typename T, int NW, int NT>
__device__ void sd_reduce(Shared<T, NW>& S, const Mem<T>& M) {
  const int tid = (int)threadIdx.x;
  const int64_t np = (int64_t)M.bstride * 8;
  const double* qn = M.sdb;
  const double* qd = M.sdb + np;
  double* ln = M.sdb + 2 * np;
  double* ld = ln + sd_leaves(np);
  const int64_t nbf = M.n / PW_BLOCK;           // full blocks
  const int64_t nfl = nbf * (PW_BLOCK / PW_LEAF);  // their leaves
  const int64_t nl = nfl + S.pw_n;
  for (int64_t j = tid; j < nl; j += NT) {
    int64_t o;
    int len;
    if (j < nfl) {
      o = j * PW_LEAF;
      len = PW_LEAF;
    } else {
      o = nbf * PW_BLOCK + S.pw_off[j - nfl];
      len = S.pw_len[j - nfl];
    }
    ln[j] = pw_leaf(qn + o, len);
    ld[j] = pw_leaf(qd + o, len);
  }
  __syncthreads();
  // full blocks: a balanced tree over their 64 leaves, in place
  for (int64_t b = tid; b < nbf; b += NT) {
    double* vn = ln + b * (PW_BLOCK / PW_LEAF);
    double* vd = ld + b * (PW_BLOCK / PW_LEAF);
    for (int st = 1; st < PW_BLOCK / PW_LEAF; st <<= 1)
      for (int k = 0; k < PW_BLOCK / PW_LEAF; k += 2 * st) {
        vn[k] = vn[k] + vn[k + st];
        vd[k] = vd[k] + vd[k + st];
      }
  }
  // the partial block: its split tree over its leaves, depth first
  if (tid == NT - 1) {
    double rn = 0.0, rd = 0.0;
    const int m = (int)(M.n % PW_BLOCK);
    if (m > 0) {
      const double* vn = ln + nfl;
      const double* vd = ld + nfl;
      int li = 0, sp = 1;
      S.pws_l[0] = m;
      S.pws_st[0] = 0;
      for (;;) {
        const int f = sp - 1;
        bool deliver = false;
        if (S.pws_st[f] == 0) {
          const int l = S.pws_l[f];
          if (l <= PW_LEAF) {
            rn = vn[li];
            rd = vd[li];
            ++li;
            --sp;
            deliver = true;
          } else {
            const int n2 = pw_split(l);
            S.pws_n2[f] = n2;
            S.pws_st[f] = 1;
            S.pws_l[sp] = n2;
            S.pws_st[sp] = 0;
            ++sp;
          }
        }
        while (deliver) {
          if (sp == 0) break;
          const int p = sp - 1;
          if (S.pws_st[p] == 1) {
            // left half done: keep it, sum the right half next
            S.pws_a[p] = rn;
            S.pws_b[p] = rd;
            S.pws_st[p] = 2;
            S.pws_l[sp] = S.pws_l[p] - S.pws_n2[p];
            S.pws_st[sp] = 0;
            ++sp;
            deliver = false;
          } else {
            rn = S.pws_a[p] + rn;
            rd = S.pws_b[p] + rd;
            --sp;
          }
        }
        if (sp == 0) break;
      }
    }
    S.pw_part[0] = rn;
    S.pw_part[1] = rd;
  }
  __syncthreads();
  if (tid == 0) {
    double nu = 0.0, de = 0.0;
    for (int64_t b = 0; b < nbf; ++b) {
      nu = nu + ln[b * (PW_BLOCK / PW_LEAF)];
      de = de + ld[b * (PW_BLOCK / PW_LEAF)];
    }
    if (M.n % PW_BLOCK) {
      nu = nu + S.pw_part[0];
      de = de + S.pw_part[1];
    }
    S.sd = (de > 0.0) ? nu / de : 0.0;
  }
}

// Positions seg_lo-1 and seg_lo of each segment need values of the previous
// segment; one thread per segment decides them after the pass and inserts
// them in front of the segment's knot list.
template <typename T, int NW>
__device__ __forceinline__ void fixup(Shared<T, NW>& S, const Mem<T>& M, const T* vals, int set, int trace) {
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
      put_knot<T>(M.kty + kix(M, set, ev, (int64_t)w * M.cap + bg), (int)q, b);
      S.beg[set][ev][w] = bg;
      S.cnt[set][ev][w] += 1;
      uint8_t* bp = M.bits + (int64_t)(set * 2 + ev) * M.bstride + seg_lo / 8;
      *bp = (uint8_t)(*bp | (1 << d));
      if (trace) atomicAdd(&S.hsh[set][ev], mix64((unsigned long long)q));
    }
  }
}

template <typename T, int NW>
__device__ __forceinline__ void prefix(Shared<T, NW>& S, int set) {
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
__device__ __forceinline__ void setup_ends(Shared<T, NW>& S, const Mem<T>& M, int cur, int end_mode) {
  // learned ends: when both envelopes have MLP_CTX extrema, learned_ends()
  // replaces the whole extension, so only the counts are set here; otherwise
  // the mirror extension stays in place
  bool counts_only = false;
  if (end_mode == END_LEARNED) {
    counts_only = S.pre[cur][0][NW] >= MLP_CTX && S.pre[cur][1][NW] >= MLP_CTX;
    end_mode = END_MIRROR;
  }
  for (int ev = 0; ev < 2; ++ev) {
    const int p = S.pre[cur][ev][NW];
    S.p[ev] = p;
    S.nl[ev] = 0;
    S.nr[ev] = 0;
    S.P[ev] = p;
    S.cval[ev] = (T)0;
    if (counts_only) continue;
    int wh = 0;
    if (p == 1) {
      const Knot<T> k = knot_at(S, M, cur, ev, 0, wh);
      S.cval[ev] = k.y;
      continue;
    }
    if (p < 2) continue;
    const Knot<T> a = knot_at(S, M, cur, ev, 0, wh);
    const Knot<T> b = knot_at(S, M, cur, ev, 1, wh);
    int wh2 = NW - 1;
    const Knot<T> c = knot_at(S, M, cur, ev, p - 2, wh2);
    const Knot<T> d = knot_at(S, M, cur, ev, p - 1, wh2);
    if (end_mode == END_MIRROR && p >= 3) {
      // emd_ref.py:92-95: reflect the 2nd / penultimate extremum past each end
      const T lt = a.t - (b.t - a.t), rt = d.t + (d.t - c.t);
      S.vt[ev][0] = lt;
      S.vt[ev][MAX_VIRT] = rt;
      const int64_t il = kix(M, cur, ev, M.vbase), ir = kix(M, cur, ev, M.vbase + MAX_VIRT);
      Pair<T> vl, vr;
      vl.x = lt;
      vl.y = b.y;
      vr.x = rt;
      vr.y = c.y;
      M.kty[il] = vl;
      M.kty[ir] = vr;
      M.km[il] = (T)0;
      M.km[ir] = (T)0;
      S.nl[ev] = 1;
      S.nr[ev] = 1;
      S.P[ev] = p + 2;
    } else {
      // natural: M = 0 at the first and last real knot
      M.km[kix(M, cur, ev, a.code)] = (T)0;
      M.km[kix(M, cur, ev, d.code)] = (T)0;
    }
  }
}

// Learned end extension (all threads of the workgroup; call after
// setup_ends, which puts the mirror extension in place when an envelope has
// fewer than MLP_CTX extrema; then this returns and the mirror stays).  For each end, the MLP_CTX extrema nearest
// to it on both envelopes give the model input:
//   gaps   d_0 = distance of the nearest extremum from the end sample,
//          d_i = distance between the i-th and (i-1)-th nearest, as d/G - 1
//          with G the mean of the 30 inter-extremum gaps of both envelopes;
//   values (y - mu) / sd over the 32 values of both envelopes;
// the sums behind G, mu and sd are xor-butterfly trees over the end's 32
// [env][i] slots (partner offsets 16, 8, 4, 2, 1).
// Per envelope the output o[0..3] gives gaps (first: from the end sample,
// G exp(o) - 1 >= 0, kept >= 2 from the nearest extremum; then G exp(o) >= 2
// apart) and o[4..7] values mu + sd o of MLP_PRED virtual extrema beyond the
// end.  The outermost virtual knot on each side carries M = 0.
template <typename T, int NW, int NT>
__device__ __forceinline__ void learned_ends(Shared<T, NW>& S, LearnedLDS& L, const Mem<T>& M,
                                             int cur, const uint16_t* w) {
  const int tid = (int)threadIdx.x;
  if (S.p[0] < MLP_CTX || S.p[1] < MLP_CTX) return;  // workgroup-uniform
  if (tid < 4 * MLP_CTX) {
    // wave 0, lane = [end][env][i]: gather, scale statistics and features.
    // The sums over an end's 32 lanes use a fixed xor-butterfly tree
    // (partner offsets 16, 8, 4, 2, 1), so every lane holds the same value.
    const int end = tid >> 5, ev = (tid >> 4) & 1, i = tid & 15;
    const int p = S.p[ev];
    int wh = end ? NW - 1 : 0;
    const int64_t slot = real_slot(S, M, cur, ev, end ? p - 1 - i : i, wh);
    const Pair<T> r = M.kty[kix(M, cur, ev, slot)];
    const double t = (double)r.x;
    const double ls = end ? (double)(M.n - 1) - t : t;
    const float yf = (float)(double)r.y;
    const double lsp = __shfl_up(ls, 1);
    const float gap = (float)(i == 0 ? ls : ls - lsp);
    float gs = (i == 0) ? 0.0f : gap, ys = yf;
#pragma unroll
    for (int k = 16; k >= 1; k >>= 1) {
      gs += __shfl_xor(gs, k);
      ys += __shfl_xor(ys, k);
    }
    const float G = gs / (float)(2 * (MLP_CTX - 1));
    const float mu = ys / (float)(2 * MLP_CTX);
    const float dv = yf - mu;
    float vs = dv * dv;
#pragma unroll
    for (int k = 16; k >= 1; k >>= 1) vs += __shfl_xor(vs, k);
    const float sd = fmaxf(sqrtf(vs / (float)(2 * MLP_CTX)), 1e-30f);
    L.x[end][ev * 32 + i] = gap / G - 1.0f;
    L.x[end][ev * 32 + 16 + i] = (yf - mu) / sd;
    if (i == 0) L.ls[end][ev][0] = ls;
    if ((tid & 31) == 0) {
      L.g[end] = G;
      L.mu[end] = mu;
      L.sd[end] = sd;
    }
  }
  __syncthreads();
  // Layers: 512 work items [end][neuron][half]; in each wave, lanes 0-31
  // run the first-half chains and lanes 32-63 the second-half chains of the
  // same 32 neurons, joined with one cross-half exchange.  The 32 outputs of
  // the last layer stay in wave 0's lanes 0-31 ([end][output]).
  float out = 0.0f;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const int nin = l == 0 ? MLP_IN : MLP_HID;
    const int nout = l == 3 ? MLP_OUT : MLP_HID;
    const int wrow = l == 0 ? MLP_LDS_W1 : l == 1 ? MLP_LDS_W2 : l == 2 ? MLP_LDS_W3 : MLP_LDS_W4;
    const int brow = l == 0 ? MLP_LDS_B1 : l == 1 ? MLP_LDS_B2 : l == 2 ? MLP_LDS_B3 : MLP_LDS_B4;
    const int items = 2 * nout * 2;
    for (int it = tid; it < ((items + WAVE - 1) / WAVE) * WAVE; it += NT) {
      const int wv = it >> 6, ln = it & 63, half = ln >> 5;
      const int q = wv * 32 + (ln & 31);  // [end][neuron]
      const int end = q / nout, jn = q - end * nout;
      const bool live = q < 2 * nout;
      float part = 0.0f;
      if (live) {
        const float* in = l == 0 ? L.x[end] : L.h[(l - 1) & 1][end];
        const uint16_t* row = w + wrow + jn * (nin + MLP_LDS_PAD) + half * (nin / 2);
        const float acc0 = half ? 0.0f : bf16_to_f32(w[brow + jn]);
        part = (l == 0) ? mlp_half_row_n<MLP_IN / 2>(row, acc0, in + half * (nin / 2))
                        : mlp_half_row_n<MLP_HID / 2>(row, acc0, in + half * (nin / 2));
      }
      const float z = part + __shfl_xor(part, 32);
      if (l < 3) {
        if (live && half == 0) L.h[l & 1][end][jn] = mlp_gelu(z);
      } else {
        out = z;
      }
    }
    if (l < 3) __syncthreads();
  }
  if (tid < WAVE) {
    // wave 0: one virtual extremum per lane [end][env][k] (lanes 0-15); the
    // gaps of its envelope and end are recomputed in order, so the positions
    // equal the sequential construction bit for bit
    const int end = (tid >> 3) & 1, ev = (tid >> 2) & 1, kk = tid & 3;
    const int ob = end * MLP_OUT + ev * 8;  // lane holding o[0] of this envelope
    float o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = __shfl(out, ob + k);
    if (tid < 2 * 2 * MLP_PRED) {
      const float G = L.g[end], mu = L.mu[end], sd = L.sd[end];
      const double dir = end ? 1.0 : -1.0;
      const double e_end = end ? (double)(M.n - 1) : 0.0;
      const double t_near = e_end - dir * L.ls[end][ev][0];
      const float d0 = fmaxf(G * mlp_exp(fminf(fmaxf(o[0], -4.0f), 4.0f)) - 1.0f, 0.0f);
      double q = e_end + dir * (double)d0;
      if (dir * (q - t_near) < 2.0) q = t_near + dir * 2.0;
      for (int k = 1; k <= kk; ++k) {
        const float gk = fmaxf(G * mlp_exp(fminf(fmaxf(o[k], -4.0f), 4.0f)), 2.0f);
        q = q + dir * (double)gk;
      }
      float ov = o[4];
      ov = kk == 1 ? o[5] : ov;
      ov = kk == 2 ? o[6] : ov;
      ov = kk == 3 ? o[7] : ov;
      // left end: slots in time order, outermost first; right end: nearest first
      const int slot = end ? MAX_VIRT + kk : MLP_PRED - 1 - kk;
      const int64_t ix = kix(M, cur, ev, M.vbase + slot);
      Pair<T> r;
      r.x = (T)q;
      r.y = (T)(mu + sd * ov);
      M.kty[ix] = r;
      M.km[ix] = (T)0;
      S.vt[ev][slot] = (T)q;
      if (kk == 0) {
        if (end) S.nr[ev] = MLP_PRED;
        else S.nl[ev] = MLP_PRED;
        S.P[ev] = S.p[ev] + 2 * MLP_PRED;
      }
    }
  }
}

template <typename T, int NW>
__device__ __forceinline__ void store_m(Shared<T, NW>& S, const Mem<T>& M, int cur, int env,
                                        int code, double val) {
  M.km[kix(M, cur, env, code)] = (T)val;
}

// Sequential reader of the knot slots of extended indices e, e+1, ... of
// (set, env): real knots advance slot by slot inside a segment run and jump
// to the next non-empty segment at its end; virtual knots come from the
// VSLOTS records.
struct KStream {
  int e;         // extended index of the next knot
  int w;         // segment of the next real knot
  int rem;       // real knots left in the current segment run
  int64_t slot;  // slot of the next real knot
};

template <typename T, int NW>
__device__ __forceinline__ void ks_real_at(KStream& ks, const Shared<T, NW>& S, const Mem<T>& M,
                                           int set, int env, int g) {
  int w = 0;
  ks.slot = real_slot(S, M, set, env, g, w);
  ks.w = w;
  ks.rem = S.cnt[set][env][w] - (g - S.pre[set][env][w]);
}

template <typename T, int NW>
__device__ __forceinline__ void ks_seek(KStream& ks, const Shared<T, NW>& S, const Mem<T>& M,
                                        int set, int env, int e) {
  ks.e = e;
  ks.rem = 0;
  const int g = e - S.nl[env];
  if (g >= 0 && g < S.p[env]) ks_real_at(ks, S, M, set, env, g);
}

template <typename T, int NW>
__device__ __forceinline__ int64_t ks_next(KStream& ks, const Shared<T, NW>& S, const Mem<T>& M,
                                           int set, int env) {
  const int nl = S.nl[env], p = S.p[env];
  const int g = ks.e - nl;
  ks.e++;
  if (g < 0) {
    if (g == -1 && p > 0) ks_real_at(ks, S, M, set, env, 0);
    return M.vbase + (g + nl);
  }
  if (g >= p) return M.vbase + MAX_VIRT + (g - p);
  const int64_t s = ks.slot;
  if (--ks.rem > 0) {
    ks.slot++;
  } else if (g + 1 < p) {
    int w = ks.w + 1;
#pragma unroll 1
    while (S.cnt[set][env][w] == 0) ++w;
    ks.w = w;
    ks.slot = (int64_t)w * M.cap + S.beg[set][env][w];
    ks.rem = S.cnt[set][env][w];
  }
  return s;
}

template <typename T>
__device__ __forceinline__ void load_ty(const T* r, double& t, double& y) {
  if constexpr (std::is_same<T, float>::value) {
    const float2 v = *reinterpret_cast<const float2*>(r);
    t = v.x;
    y = v.y;
  } else {
    const double2 v = *reinterpret_cast<const double2*>(r);
    t = v.x;
    y = v.y;
  }
}

// Reverse knot stream: slots of extended indices e, e-1, ... (the back
// substitution walks the window from its last row to its first).
struct KRev {
  int e;         // extended index of the next knot to return
  int w;         // segment of the next real knot
  int left;      // real knots before `slot` in its segment run
  int64_t slot;  // slot of the next real knot
};

template <typename T, int NW>
__device__ __forceinline__ void kr_real_at(KRev& kr, const Shared<T, NW>& S, const Mem<T>& M,
                                           int set, int env, int g) {
  int w = NW - 1;
  kr.slot = real_slot(S, M, set, env, g, w);
  kr.w = w;
  kr.left = g - S.pre[set][env][w];
}

template <typename T, int NW>
__device__ __forceinline__ void kr_seek(KRev& kr, const Shared<T, NW>& S, const Mem<T>& M, int set,
                                        int env, int e) {
  kr.e = e;
  kr.left = 0;
  const int g = e - S.nl[env];
  if (g >= 0 && g < S.p[env]) kr_real_at(kr, S, M, set, env, g);
}

template <typename T, int NW>
__device__ __forceinline__ int64_t kr_prev(KRev& kr, const Shared<T, NW>& S, const Mem<T>& M,
                                           int set, int env, int nl, int p) {
  const int g = kr.e - nl;
  kr.e--;
  if (g >= p) {
    if (g == p && p > 0) kr_real_at(kr, S, M, set, env, p - 1);
    return M.vbase + MAX_VIRT + (g - p);
  }
  if (g < 0) return M.vbase + (g + nl);
  const int64_t s = kr.slot;
  if (kr.left > 0) {
    kr.slot--;
    kr.left--;
  } else if (g >= 1) {
    int w = kr.w - 1;
#pragma unroll 1
    while (S.cnt[set][env][w] == 0) --w;
    kr.w = w;
    kr.left = S.cnt[set][env][w] - 1;
    kr.slot = (int64_t)w * M.cap + S.beg[set][env][w] + kr.left;
  }
  return s;
}

// Division for the window recurrences: IEEE division in FP64 mode (the
// oracle's rounding, emd_ref.py:58-61); in FP32 mode one FP64 reciprocal
// per denominator.
template <typename T>
__device__ __forceinline__ double qdiv(double a, double b) {
  if constexpr (std::is_same<T, double>::value) return a / b;
  else return a * __builtin_amdgcn_rcp(b);
}

// One window [w0, w1] of rows (couplings outside the window dropped) solved
// by a whole wave with the partition method: lane L eliminates its block of
// sb rows downwards keeping the dependence on the row before the block
// (x_i + c'_i x_{i+1} = d'_i + v_i alpha), sweeps back to express its first
// and last unknowns in (alpha, beta), the wave solves the 2-per-lane
// interface system serially with readlane broadcasts, and every lane
// reconstructs its rows.  Exact up to rounding (FP32 mode only).
template <typename T, int NW>
__device__ __noinline__ void window_wave(Shared<T, NW>& S, const Mem<T>& M, int cur, int ev,
                                            int n, int w0, int w1, int o0, int o1) {
  const int ln = lane_id();
  const int w = (int)(threadIdx.x / WAVE);
  const int m = w1 - w0 + 1;
  const int sb = (m + WAVE - 1) / WAVE;
  const int nact = (m + sb - 1) / sb;
  const int s = w0 + ln * sb;
  const int e = (s + sb - 1 < w1) ? s + sb - 1 : w1;
  const bool act = ln < nact;
  const Pair<T>* const tyb = M.kty + kix(M, cur, ev, 0);
  T* const mb = M.km + kix(M, cur, ev, 0);
  float* const scr = reinterpret_cast<float*>(M.wcx) + (int64_t)w * (3 * 64 * WAVE);
  // scratch row j of this lane: scr[(j*3 + comp) * WAVE + ln], sb <= 64
  double Ps = 0, Qs = 0, Rs = 0, Pe = 0, Qe = 0, Re = 0;
  if (act) {
    KStream ks;
    ks_seek(ks, S, M, cur, ev, s);
    float2 ka = tyb[ks_next(ks, S, M, cur, ev)];
    float2 kb = tyb[ks_next(ks, S, M, cur, ev)];
    double at = ka.x, ay = ka.y, bt = kb.x, by = kb.y;
    double sl0 = (by - ay) * __builtin_amdgcn_rcp(bt - at);
    double cp = 0.0, dp = 0.0, vp = 0.0;
#pragma unroll 1
    for (int q = s; q <= e; ++q) {
      const float2 kc = tyb[ks_next(ks, S, M, cur, ev)];
      const double ct = kc.x, cy = kc.y;
      const double H0 = bt - at, H1 = ct - bt;
      const double d = 2.0 * (H0 + H1);
      const double l = (q >= 1 && q > w0) ? H0 : 0.0;   // window edge: coupling dropped
      const double u = (q <= n - 2 && q < w1) ? H1 : 0.0;
      const double sl1 = (cy - by) * __builtin_amdgcn_rcp(H1);
      const double r = 6.0 * (sl1 - sl0);
      const double den = (q == s) ? d : d - l * cp;
      const double inv = __builtin_amdgcn_rcp(den);
      const double c = u * inv;
      const double dd = ((q == s) ? r : r - l * dp) * inv;
      const double vv = ((q == s) ? -l : -l * vp) * inv;
      const int j = q - s;
      scr[(j * 3 + 0) * WAVE + ln] = (float)c;
      scr[(j * 3 + 1) * WAVE + ln] = (float)dd;
      scr[(j * 3 + 2) * WAVE + ln] = (float)vv;
      cp = c;
      dp = dd;
      vp = vv;
      sl0 = sl1;
      at = bt;
      ay = by;
      bt = ct;
      by = cy;
    }
    // last row: x_e = d'_e + v_e alpha - c'_e beta
    Pe = dp;
    Qe = vp;
    Re = -cp;
    double P = Pe, Q = Qe, R = Re;
#pragma unroll 1
    for (int j = e - s - 1; j >= 0; --j) {
      const double c = scr[(j * 3 + 0) * WAVE + ln];
      const double dd = scr[(j * 3 + 1) * WAVE + ln];
      const double vv = scr[(j * 3 + 2) * WAVE + ln];
      P = dd - c * P;
      Q = vv - c * Q;
      R = -c * R;
    }
    Ps = P;
    Qs = Q;
    Rs = R;
  }
  // interface system: F_L = x_{s_L}, G_L = x_{e_L}; G_{-1} = F_{nact} = 0
  //   F_L = Ps + Qs G_{L-1} + Rs F_{L+1},  G_L = Pe + Qe G_{L-1} + Re F_{L+1}
  // forward: G_{L-1} = A + B F_L  ->  F_L = a_L + b_L F_{L+1},  G_L = A_L + B_L F_{L+1}
  double A = 0.0, B = 0.0;
  double my_a = 0.0, my_b = 0.0, my_A = 0.0, my_B = 0.0;
#pragma unroll 1
  for (int L = 0; L < nact; ++L) {
    // L is wave-uniform: v_readlane instead of an LDS permute per value
    const double ps = lane_val(Ps, L), qs = lane_val(Qs, L), rs = lane_val(Rs, L);
    const double pe = lane_val(Pe, L), qe = lane_val(Qe, L), re = lane_val(Re, L);
    const double inv = __builtin_amdgcn_rcp(1.0 - qs * B);
    const double a = (ps + qs * A) * inv;
    const double b = rs * inv;
    const double An = pe + qe * (A + B * a);
    const double Bn = qe * B * b + re;
    if (ln == L) {
      my_a = a;
      my_b = b;
      my_A = An;
      my_B = Bn;
    }
    A = An;
    B = Bn;
  }
  double alpha = 0.0, beta = 0.0, Fn = 0.0;
#pragma unroll 1
  for (int L = nact - 1; L >= 0; --L) {
    const double fL = lane_val(my_a + my_b * Fn, L);
    const double gL = lane_val(my_A + my_B * Fn, L);
    if (ln == L) beta = Fn;
    if (ln == L + 1) alpha = gL;
    Fn = fL;
  }
  if (act) {
    // reconstruct rows e .. s (row q sets M of extended knot q+1)
    const int nl = S.nl[ev], p = S.p[ev];
    KRev kr;
    kr_seek(kr, S, M, cur, ev, e + 1);
    double x = Pe + Qe * alpha + Re * beta;
    {
      const int64_t sl = kr_prev(kr, S, M, cur, ev, nl, p);
      if (e >= o0 && e <= o1) mb[sl] = (T)x;
    }
#pragma unroll 1
    for (int j = e - s - 1; j >= 0; --j) {
      const double c = scr[(j * 3 + 0) * WAVE + ln];
      const double dd = scr[(j * 3 + 1) * WAVE + ln];
      const double vv = scr[(j * 3 + 2) * WAVE + ln];
      x = dd + vv * alpha - c * x;
      const int64_t sl = kr_prev(kr, S, M, cur, ev, nl, p);
      const int q = s + j;
      if (q >= o0 && q <= o1) mb[sl] = (T)x;
    }
  }
}

// Decay-seamed block Thomas over both envelopes of set `cur`: owned blocks of
// B rows solved on the clipped window [o0-S, o1+S] with the couplings outside
// the window dropped (emd_gpu_v5_dsbt.py:51-74, 223-237); one window per
// thread.  B <= 0: one window per envelope (exact Thomas, emd_ref.py:50-65).
// Rows are built on the fly from the knots (emd_ref.py:31-48).  The serial
// recurrence is fed by a knot stream read PF rows ahead; the back
// substitution reads its forward-sweep rows PF ahead and finds each row's
// knot slot by walking the knots backwards.
template <typename T, int NW>
__device__ __forceinline__ void solve(Shared<T, NW>& S, const Mem<T>& M, int cur, int Bp, int Sp,
                                      int wave_windows) {
  constexpr int PF = 16;
  using T2 = typename std::conditional<std::is_same<T, float>::value, float2, double2>::type;
  T2* const wcx = reinterpret_cast<T2*>(M.wcx);
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
  if constexpr (std::is_same<T, float>::value) {
    // few long windows: a wave per window (partition method)
    // needs blocks longer than a wave, windows of at most 64 rows per lane
    // and the per-wave scratch (3 x 64 x 64 floats) inside the window area
    const bool fits = M.cols >= (int64_t)blockDim.x &&
                      M.rows * M.cols * 2 >= (int64_t)NW * 3 * 64 * WAVE &&
                      Be[0] + 2 * Sp <= 64 * WAVE && Be[1] + 2 * Sp <= 64 * WAVE;
    if (fits && total <= wave_windows && Be[0] + Be[1] > 2 * WAVE) {
      const int wv = (int)(threadIdx.x / WAVE);
#pragma unroll 1
      for (int wi = wv; wi < total; wi += NW) {
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
        window_wave(S, M, cur, ev, n, w0, w1, o0, o1);
      }
      return;
    }
  }
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
    const Pair<T>* const tyb = M.kty + kix(M, cur, ev, 0);
    T* const mb = M.km + kix(M, cur, ev, 0);

    KStream ks;
    ks_seek(ks, S, M, cur, ev, w0);
    const T2 ka = tyb[ks_next(ks, S, M, cur, ev)];
    const T2 kb = tyb[ks_next(ks, S, M, cur, ev)];
    double at = (double)ka.x, ay = (double)ka.y, bt = (double)kb.x, by = (double)kb.y;
    double sl0 = qdiv<T>(by - ay, bt - at);
    double cprev = 0.0, xprev = 0.0;

    const int elast = w1 + 2;  // last knot the window's rows need
    int enext = w0 + 2;
    // prefetched knots stay in their storage type until used, so the loads
    // are not waited for at issue
    T2 nk[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      if (enext <= elast) {
        nk[u] = tyb[ks_next(ks, S, M, cur, ev)];
        ++enext;
      } else {
        nk[u].x = (T)0;
        nk[u].y = (T)0;
      }
    }
    const int nl_ev = S.nl[ev], p_ev = S.p[ev];
    // one row of the forward sweep (emd_gpu_v5_dsbt.py:58-66, the system of
    // emd_ref.py:31-48 built from the knots)
    auto row = [&](int q, T2 kn) {
      const double ct = (double)kn.x, cy = (double)kn.y;
      const double H0 = bt - at;
      const double H1 = ct - bt;
      const double d = 2.0 * (H0 + H1);
      const double l = (q >= 1) ? H0 : 0.0;
      // no upper coupling on the window's last row (the reference's 0.0) or
      // on the system's last row: a zero numerator gives that zero (den > 0),
      // so the division needs no branch
      const double uu = (q <= n - 2 && q != w1) ? H1 : 0.0;
      double sl1, c, xv;
      const double den = d - l * cprev;
      if constexpr (std::is_same<T, double>::value) {
        sl1 = (cy - by) / H1;
        const double r = 6.0 * (sl1 - sl0);
        c = uu / den;
        xv = (r - l * xprev) / den;
      } else {
        sl1 = qdiv<T>(cy - by, H1);
        const double r = 6.0 * (sl1 - sl0);
        const double inv = __builtin_amdgcn_rcp(den);
        c = uu * inv;
        xv = (r - l * xprev) * inv;
      }
      T2 cx;
      cx.x = (T)c;
      cx.y = (T)xv;
      wcx[(int64_t)(q - w0) * M.cols + col] = cx;
      cprev = c;
      xprev = xv;
      sl0 = sl1;
      at = bt;
      ay = by;
      bt = ct;
      by = cy;
    };
#pragma unroll 1
    for (int q0 = w0; q0 <= w1; q0 += PF) {
      T2 fk[PF];
      // the next PF knots are consecutive slots of one real run inside the
      // window: read them directly (FP32: 2 pairs per 16-byte load, half the
      // requests of per-pair loads; each lane streams its own window, so
      // these are uncoalesced and request-bound)
      const int g = ks.e - nl_ev;
      const bool wide = g >= 0 && g + PF < p_ev && ks.rem > PF && enext + PF - 1 <= elast;
      if (wide) {
        const T2* src = tyb + ks.slot;
        if constexpr (std::is_same<T, float>::value) {
          if ((ks.slot & 1) == 0) {
#pragma unroll
            for (int u = 0; u < PF; u += 2) {
              const float4 v = *reinterpret_cast<const float4*>(src + u);
              fk[u].x = v.x;
              fk[u].y = v.y;
              fk[u + 1].x = v.z;
              fk[u + 1].y = v.w;
            }
          } else {
            fk[0] = src[0];
#pragma unroll
            for (int u = 1; u < PF - 1; u += 2) {
              const float4 v = *reinterpret_cast<const float4*>(src + u);
              fk[u].x = v.x;
              fk[u].y = v.y;
              fk[u + 1].x = v.z;
              fk[u + 1].y = v.w;
            }
            fk[PF - 1] = src[PF - 1];
          }
        } else {
#pragma unroll
          for (int u = 0; u < PF; ++u) fk[u] = src[u];
        }
        ks.slot += PF;
        ks.rem -= PF;
        ks.e += PF;
        enext += PF;
      } else {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          if (enext <= elast) {
            fk[u] = tyb[ks_next(ks, S, M, cur, ev)];
            ++enext;
          } else {
            fk[u].x = (T)0;
            fk[u].y = (T)0;
          }
        }
      }
      if (q0 + PF - 1 <= w1) {
        // a whole group of rows: straight-line code, so the work of a row
        // (and the slopes of later rows) overlaps
#pragma unroll
        for (int u = 0; u < PF; ++u) row(q0 + u, nk[u]);
      } else {
#pragma unroll
        for (int u = 0; u < PF; ++u)
          if (q0 + u <= w1) row(q0 + u, nk[u]);
      }
#pragma unroll
      for (int u = 0; u < PF; ++u) nk[u] = fk[u];
    }
    // back substitution, rows w1 .. w0 (row q sets M of extended knot q+1)
    const int nl = S.nl[ev], p = S.p[ev];
    KRev kr;
    kr_seek(kr, S, M, cur, ev, w1 + 1);
    double xn = xprev;
    // owned M values of adjacent slots (s even, s + 1) leave as one paired
    // store; an odd slot waits one row for its even neighbour
    int64_t pend_s = -1;
    T pend_v = (T)0;
    auto put_m = [&](int64_t sq, T v) {
      if (pend_s == sq + 1 && (sq & 1) == 0) {
        T2 mv;
        mv.x = v;
        mv.y = pend_v;
        *reinterpret_cast<T2*>(mb + sq) = mv;
        pend_s = -1;
        return;
      }
      if (pend_s >= 0) mb[pend_s] = pend_v;
      pend_s = -1;
      if (sq & 1) {
        pend_s = sq;
        pend_v = v;
      } else {
        mb[sq] = v;
      }
    };
    {
      const int64_t s1 = kr_prev(kr, S, M, cur, ev, nl, p);
      if (w1 >= o0 && w1 <= o1) put_m(s1, (T)xn);
    }
    T2 cx[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int q = w1 - 1 - u;
      cx[u] = wcx[(int64_t)((q < w0 ? w0 : q) - w0) * M.cols + col];
    }
#pragma unroll 1
    for (int q0 = w1 - 1; q0 >= w0; q0 -= PF) {
      T2 fx[PF];
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int q = q0 - PF - u;
        fx[u] = wcx[(int64_t)((q < w0 ? w0 : q) - w0) * M.cols + col];
      }
      // a whole group of owned rows whose knots are consecutive slots of one
      // real run (the window's interior): straight-line rows, direct slots
      const int g = kr.e - nl;
      if (q0 - PF + 1 >= w0 && q0 - PF + 1 >= o0 && q0 <= o1 && g < p && kr.left >= PF) {
        if (pend_s >= 0) mb[pend_s] = pend_v;
        pend_s = -1;
        T* const ms = mb + kr.slot;
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const double xi = (double)cx[u].y - (double)cx[u].x * xn;
          ms[-u] = (T)xi;
          xn = xi;
        }
        kr.slot -= PF;
        kr.left -= PF;
        kr.e -= PF;
      } else {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int q = q0 - u;
          if (q >= w0) {
            const double xi = (double)cx[u].y - (double)cx[u].x * xn;
            const int64_t sq = kr_prev(kr, S, M, cur, ev, nl, p);
            if (q >= o0 && q <= o1) put_m(sq, (T)xi);
            xn = xi;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < PF; ++u) cx[u] = fx[u];
    }
    if (pend_s >= 0) mb[pend_s] = pend_v;
  }
}

// ------------------------------------------------------------------------
// the persistent kernel
// ------------------------------------------------------------------------
__device__ __forceinline__ uint64_t clk() { return wall_clock64(); }

// Waves per SIMD the kernel is register-limited to.  A 64-thread workgroup
// (one wave, one member at a time) in FP32 keeps two waves per SIMD, as the
// 512-thread workgroup must; larger workgroups get what their size implies.
template <typename T, int NT>
constexpr int sift_min_waves() { return (NT == 64 && sizeof(T) == 4) ? 2 : 1; }

template <typename T, int NT>
__global__ __launch_bounds__(NT, (sift_min_waves<T, NT>())) void sift_kernel(SiftArgs A) {
  constexpr int NW = NT / WAVE;
  __shared__ Shared<T, NW> S;
  __shared__ LearnedLDS LL;
  extern __shared__ __align__(16) uint16_t model_lds[];  // learned mode: MLP_LDS_BYTES
  const int tid = (int)threadIdx.x;

  char* scr = A.scratch + (int64_t)blockIdx.x * A.scratch_stride;
  const ScratchLayout LY = make_layout((int)sizeof(T), A.npad, NW, A.cap, A.rows, A.cols,
                                       A.tau > 0.0 ? 1 : 0);
  Mem<T> M;
  M.cap = A.cap;
  M.h = reinterpret_cast<T*>(scr + LY.h);
  M.r = reinterpret_cast<T*>(scr + LY.r);
  M.bits = reinterpret_cast<uint8_t*>(scr + LY.bits);
  M.bstride = A.npad / 8;
  M.kty = reinterpret_cast<Pair<T>*>(scr + LY.kty);
  M.km = reinterpret_cast<T*>(scr + LY.km);
  M.rstride = (int64_t)NW * A.cap + VSLOTS;
  M.vbase = (int64_t)NW * A.cap;
  M.wcx = reinterpret_cast<void*>(scr + LY.wcx);
  M.sdb = reinterpret_cast<double*>(scr + LY.sd);
  M.rows = A.rows;
  M.cols = A.cols;

  const int trace = A.trace != nullptr;
  const bool timing = A.phase != nullptr;
  uint64_t ph[4] = {0, 0, 0, 0};
  uint64_t tc = timing ? clk() : 0;
#define PHASE(k)                          \
  if (timing && tid == 0) {               \
    const uint64_t t_ = clk();            \
    ph[k] += t_ - tc;                     \
    tc = t_;                              \
  }

  // The BF16 model, one row per neuron (the layout of MLP_LDS_*): a copy in
  // LDS for the workgroup's lifetime, except for one-wave workgroups, which
  // read the same image from global memory (L2-resident) so that many of them
  // fit on a CU.
  const uint16_t* model = model_lds;
  if (A.end_mode == END_LEARNED) {
    if constexpr (NT == WAVE) {
      model = reinterpret_cast<const uint16_t*>(A.mlp_img);
    } else {
      // global layout W[in][out] read coalesced
      const uint16_t* g = reinterpret_cast<const uint16_t*>(A.mlp);
      const int gw[4] = {MLP_W1, MLP_W2, MLP_W3, MLP_W4};
      const int gb[4] = {MLP_B1, MLP_B2, MLP_B3, MLP_B4};
      const int lw[4] = {MLP_LDS_W1, MLP_LDS_W2, MLP_LDS_W3, MLP_LDS_W4};
      const int lb[4] = {MLP_LDS_B1, MLP_LDS_B2, MLP_LDS_B3, MLP_LDS_B4};
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        const int nin = l == 0 ? MLP_IN : MLP_HID, nout = l == 3 ? MLP_OUT : MLP_HID;
        for (int k = tid; k < nin * nout; k += NT) {
          const int i = k / nout, j = k - i * nout;
          model_lds[lw[l] + j * (nin + MLP_LDS_PAD) + i] = g[gw[l] + k];
        }
        for (int j = tid; j < nout; j += NT) model_lds[lb[l] + j] = g[gb[l] + j];
      }
    }
  }

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
      io.sig = (uint32_t)(A.signal0 + td.signal);
      io.mem = (uint32_t)(A.member0 + mem);
      io.trace = trace;
      __syncthreads();
      if (tid == 0) {
        S.hsh[0][0] = S.hsh[0][1] = 0ull;
        S.hsh[1][0] = S.hsh[1][1] = 0ull;
        if (A.tau > 0.0) pw_enumerate(S, M.n);
      }
      __syncthreads();
      PHASE(3);
      wave_pass<T, NW, PASS_INIT>(S, M, io, 1, 0);
      __syncthreads();
      PHASE(1);
      fixup(S, M, M.r, 0, trace);
      __syncthreads();
      if (tid == 0) prefix(S, 0);
      __syncthreads();
      PHASE(2);

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
          if (A.end_mode == END_LEARNED) {
            learned_ends<T, NW, NT>(S, LL, M, cur, model);
            __syncthreads();
          }
          PHASE(3);
          solve(S, M, cur, A.B, A.S, A.wave_windows);
          __syncthreads();
          PHASE(0);
          const bool last = (s == A.max_sifts - 1);
          io.src_is_r = (s == 0);
          if (last) {
            io.imf = A.member_imfs
                         ? (void*)(reinterpret_cast<T*>(A.member_imfs) + sd.mimf_off +
                                   ((int64_t)mem * A.max_imf + m) * sd.n)
                         : nullptr;
            io.acc = A.acc ? A.acc + td.acc_off + (int64_t)m * sd.n : nullptr;
            io.acc_first = (S.acc_cnt[m] == 0);
            wave_sift_pipe<T, NW, true>(S, M, io, cur, nxt);
          } else {
            io.want_sd = (A.tau > 0.0) && (s >= 1);
            wave_sift_pipe<T, NW, false>(S, M, io, cur, nxt);
          }
          __syncthreads();
          PHASE(1);
          fixup(S, M, last ? M.r : M.h, nxt, trace);
          __syncthreads();
          if (tid == 0) prefix(S, nxt);
          if (!last && io.want_sd) sd_reduce<T, NW, NT>(S, M);
          __syncthreads();
          PHASE(2);
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
          PHASE(3);
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
  if (timing && tid == 0) {
    PHASE(3);
    for (int k = 0; k < 4; ++k) A.phase[(int64_t)blockIdx.x * 4 + k] = ph[k];
  }
#undef PHASE
}

// ------------------------------------------------------------------------
// host-side launchers (called from the C-ABI)
// ------------------------------------------------------------------------
// Dynamic LDS of a launch: the BF16 boundary model in learned mode.
static inline size_t sift_dyn_lds(int end_mode, int nt) {
  return (end_mode == END_LEARNED && nt != WAVE) ? MLP_LDS_BYTES : 0;
}

template <typename T, int NT>
static hipError_t allow_model_lds() {
  static const hipError_t e = hipFuncSetAttribute(
      reinterpret_cast<const void*>(sift_kernel<T, NT>),
      hipFuncAttributeMaxDynamicSharedMemorySize, (int)MLP_LDS_BYTES);
  return e;
}

template <typename T, int NT>
static hipError_t launch_sift_t(const SiftArgs& A, int grid, hipStream_t st) {
  const size_t lds = sift_dyn_lds(A.end_mode, NT);
  if (lds) {
    const hipError_t e = allow_model_lds<T, NT>();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((sift_kernel<T, NT>), dim3(grid), dim3(NT), lds, st, A);
  return hipGetLastError();
}

// Occupancy (workgroups per CU) of one sift kernel instantiation.
template <typename T, int NT>
static hipError_t sift_occupancy_t(int end_mode, int* blocks_per_cu) {
  const size_t lds = sift_dyn_lds(end_mode, NT);
  if (lds) {
    const hipError_t e = allow_model_lds<T, NT>();
    if (e != hipSuccess) return e;
  }
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocks_per_cu, reinterpret_cast<const void*>(sift_kernel<T, NT>), NT, lds);
}

}  // namespace eemd
