// Shared host/device definitions of the MI355X EEMD engine.
//
// Work decomposition (one persistent launch):
//   unit  = (signal, ensemble member): one full EMD (all IMFs, all sifts).
//   task  = (signal, contiguous member range): the queue entry one
//           workgroup pops; its members are decomposed one after another
//           and their IMFs are summed, in member order, into the task's
//           FP64 accumulator (deterministic ensemble mean).
//   wave  = one contiguous segment of the member's signal; each wave streams
//           its segment in chunks of 512 samples (8 consecutive samples per
//           lane), so no barrier is needed inside a pass.
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define __host__
#define __device__
#endif

typedef struct eemd_ctx eemd_ctx;

namespace eemd {

constexpr int WAVE = 64;
constexpr int RPL = 8;               // samples per lane per chunk
constexpr int CHUNK = WAVE * RPL;    // samples per wave per chunk
constexpr int SEG_FRONT = 2;         // knot slots reserved in front of a segment region
constexpr int MAX_VIRT = 4;          // virtual (extension) knots per signal end
constexpr int MAX_IMF_CAP = 64;      // hard cap of IMFs per member (emd_par.py:205-206)

enum EndMode : int32_t { END_NATURAL = 0, END_MIRROR = 1, END_LEARNED = 2 };

struct SignalDesc {
  int64_t x_off;      // first sample of this signal in the concatenated input
  int64_t n;          // length
  int64_t noise_off;  // explicit standard normals [E][n] at this offset; <0: Philox
  int64_t avg_off;    // ensemble-mean output block [max_imf][n]
  int64_t mimf_off;   // per-member IMF block [E][max_imf][n]; <0: not stored
  double sigma;       // noise scale eps*std(x)
  int32_t unit0;      // index of (this signal, member 0) in per-unit outputs
  int32_t task0;      // first task of this signal
  int32_t ntask;      // number of tasks of this signal
  int32_t nmember;    // ensemble size E
};

struct TaskDesc {
  int32_t signal;
  int32_t m0, m1;     // members [m0, m1)
  int32_t pad;
  int64_t acc_off;    // FP64 accumulator block [max_imf][n]
};

// Per-workgroup scratch (global memory).  Offsets in bytes, 256-aligned.
// Knot records are {t, y, M, -} in the working type: 16 bytes in FP32 mode,
// 32 bytes in FP64 mode.
constexpr int VSLOTS = 2 * MAX_VIRT;  // virtual-knot records after the segment regions

struct ScratchLayout {
  int64_t h, r;       // working h and residual (T[npad])
  int64_t bits;       // extrema bitmaps uint8[2 sets][2 env][npad/8]
  int64_t kty;        // knot (position, value) pairs [2 sets][2 env][nseg*cap + VSLOTS] x 2 T
  int64_t km;         // knot second derivatives, same slots, T each
  int64_t wcx;        // window Thomas forward sweep {c', x'} [rows][cols]
                      // (double2 in FP64 mode, float2 in FP32 mode)
  int64_t sd;         // SD stop terms (tau > 0 only): per sample (h - h')^2 and
                      // h^2 [2][npad] double, then their leaf sums [2][sd_leaves]
  int64_t total;
};

// numpy's pairwise summation (np.sum of a contiguous float64 array): blocks
// of PW_BLOCK elements added in sequence; inside a block, ranges longer than
// PW_LEAF split in two; a range of at most PW_LEAF elements is a leaf.
constexpr int PW_BLOCK = 8192;
constexpr int PW_LEAF = 128;
constexpr int PW_MAXPART = 130;  // leaves of a partial block (each >= 64 elements)

// Leaf-sum slots of the SD region for a padded length npad.
__host__ __device__ inline int64_t sd_leaves(int64_t npad) { return npad / PW_LEAF + PW_MAXPART + 2; }

__host__ __device__ inline int64_t align256(int64_t v) { return (v + 255) & ~int64_t(255); }

__host__ __device__ inline ScratchLayout make_layout(int tbytes, int64_t npad, int nseg,
                                                     int64_t cap, int64_t rows, int64_t cols,
                                                     int sd) {
  ScratchLayout L;
  int64_t o = 0;
  L.h = o;    o = align256(o + npad * tbytes);
  L.r = o;    o = align256(o + npad * tbytes);
  L.bits = o; o = align256(o + 4 * (npad / 8));
  L.kty = o;  o = align256(o + 4 * (nseg * cap + VSLOTS) * 2 * tbytes);
  L.km = o;   o = align256(o + 4 * (nseg * cap + VSLOTS) * tbytes);
  L.wcx = o;  o = align256(o + rows * cols * 2 * tbytes);
  L.sd = o;   if (sd) o = align256(o + 8 * (2 * npad + 2 * sd_leaves(npad)));
  L.total = o;
  return L;
}

// Segment length for a member of length n split over nseg waves: a multiple
// of CHUNK so every chunk lies inside one segment.
__host__ __device__ inline int64_t seg_len(int64_t n, int nseg) {
  int64_t per = (n + nseg - 1) / nseg;
  return ((per + CHUNK - 1) / CHUNK) * CHUNK;
}

// Knot slots per segment region: strict extrema are >= 2 samples apart, so a
// segment of L samples holds at most L/2 + 1, plus the fix-up front slots.
__host__ __device__ inline int64_t seg_cap(int64_t seglen) { return seglen / 2 + 2 + SEG_FRONT; }

struct SiftArgs {
  const double* x;          // concatenated signals
  const double* noise;      // explicit standard normals (or null)
  const SignalDesc* sig;
  const TaskDesc* task;
  const int32_t* order;     // queue position -> task (longest first)
  int32_t ntask;
  int32_t nsig;
  int32_t* queue;           // work counter, zeroed before launch
  uint64_t seed;
  int32_t max_imf, max_sifts, min_extrema, end_mode;
  double tau;
  int32_t B, S;             // DSBT block / halo (B <= 0: exact Thomas)
  double* acc;              // FP64 task accumulators
  int32_t* acc_cnt;         // [ntask][max_imf] members that produced each level
  void* member_imfs;        // T per-member IMFs (optional)
  int32_t* nmode;           // [units]
  int32_t* nsift;           // [units]
  int32_t* per_mode;        // [units][max_imf]
  uint64_t* trace;          // [units][max_imf][max_sifts][4] (optional)
  char* scratch;
  int64_t scratch_stride;   // bytes per workgroup
  int64_t npad;             // padded max length
  int64_t cap;              // knot slots per segment region
  int64_t rows, cols;       // window scratch shape
  const void* mlp;          // learned-boundary model (optional), BF16 W[in][out] per layer
  const void* mlp_img;      // the same model as one row per neuron (MLP_LDS_* layout)
  uint64_t* phase;          // optional per-workgroup phase clocks [nslots][4]
  int32_t wave_windows;     // FP32: wave-parallel window solve when windows <= this
  int32_t nslots;           // number of workgroups (scratch slots)
  int32_t member0;          // Philox member index of the first member
  int32_t signal0;          // Philox signal index of the first signal
};

// Context accessors for the other translation units (defined with the
// context in eemd_capi.cpp): its device, its stream (a hipStream_t), its
// error message, and a slot for state another unit keeps per context, freed
// by the given function when the context is destroyed.
int ctx_device(const eemd_ctx* c);
void* ctx_stream(eemd_ctx* c);
int ctx_fail(eemd_ctx* c, int code, const char* msg);
void*& ctx_spectra(eemd_ctx* c);
void ctx_spectra_dtor(eemd_ctx* c, void (*f)(void*));

}  // namespace eemd
