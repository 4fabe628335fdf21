// C ABI of the MI355X EEMD engine (declared in include/eemd_mi355x.h).
//
// Host-side responsibilities: validate arguments, lay out the work (signal and
// task descriptors, queue order longest-first), size and reuse device scratch
// owned by the context, enqueue fill -> persistent sift -> mean kernels, and
// time the sift and mean kernels with HIP events on the launch stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/eemd_mi355x.h"
#include "eemd_common.h"
#include "eemd_mlp.h"

namespace eemd {
hipError_t launch_sift(const SiftArgs& A, int fp64, int nt, int grid, hipStream_t st);
hipError_t sift_occupancy(int fp64, int nt, int end_mode, int* blocks_per_cu);
hipError_t launch_mean(const SignalDesc* sig, int nsig, const int64_t* sig_start,
                       const TaskDesc* task, const double* acc, const int32_t* acc_cnt,
                       int max_imf, const double* x, double* avg, double* resid,
                       int32_t* nmodes, int64_t total_n, int sums, hipStream_t st);
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);
hipError_t launch_fill_f64(double* p, int64_t n, double v, hipStream_t st);
hipError_t launch_fill_bytes(void* p, int64_t n, hipStream_t st);
hipError_t launch_noise(uint64_t seed, uint32_t sig, uint32_t m0, int E, int64_t n, double* out,
                        hipStream_t st);
}  // namespace eemd

using namespace eemd;

// Device memory of the allocation-free entry point (eemd_run_batch_dev_noalloc):
// every GPU the library runs on holds its own copy until the process ends.
// Host code never uses this name as an address: the address on each device
// comes from hipGetSymbolAddress at context creation (noalloc_addr below).
static __device__ __align__(256) unsigned char scratch[EEMD_NOALLOC_SCRATCH_BYTES];

namespace {

// Per-device address of `scratch`, looked up once per device with that
// device current.
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[64] = {};
hipError_t g_noalloc_err[64] = {};
bool g_noalloc_done[64] = {};

void noalloc_lookup(int device) {
  if (device < 0 || device >= 64) return;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_done[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) (void)hipGetLastError();
  g_noalloc_ptr[device] = (e == hipSuccess) ? ptr : nullptr;
  g_noalloc_err[device] = e;
  g_noalloc_done[device] = true;
}

// A device buffer that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t want) {
    if (want <= bytes && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      bytes = 0;
      if (e != hipSuccess) return e;
    }
    size_t b = std::max<size_t>(want, 256);
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    bytes = b;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

}  // namespace

// Signal / task descriptors, queue order and signal starts of one launch.
struct DescBufs {
  DevBuf sigd, taskd, order, sig_start;
  void release() {
    sigd.release();
    taskd.release();
    order.release();
    sig_start.release();
  }
};

// One instantiated launch sequence (clears -> sift -> mean) and the exact
// call it was captured for.  It owns its descriptors, so eager calls in
// between do not disturb it.
struct GraphCache {
  hipGraphExec_t exec = nullptr;
  std::vector<uint8_t> key;
  DescBufs desc;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
  }
};

// Launch configuration chosen by eemd_autotune for one batch shape.
struct TuneEntry {
  std::vector<int64_t> key;
  int threads, mpt, wave_windows, block;
};

struct eemd_ctx {
  int device = 0;
  int ncu = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  // recorded after every call's device work: the next call's stream waits
  // for it, so calls enqueued on different streams never share the
  // context's buffers while a previous kernel still uses them (growing a
  // buffer is safe on its own: hipFree synchronizes the device)
  hipEvent_t done = nullptr;
  bool done_recorded = false;
  std::string err;
  // device buffers owned by the context
  DevBuf scratch, acc, acc_cnt, queue, phase;
  DescBufs desc;
  // staging for the host-pointer entry point
  DevBuf hx, hnoise, havg, hres, hnm, hnmode, hnsift, hpm, hmi, htr;
  // outputs of the autotuner's timed runs
  DevBuf tx, tavg, tres, tnm, tnmode, tnsift, tpm;
  DevBuf mlpw;  // BF16 boundary model (EEMD_END_LEARNED)
  DevBuf mlpi;  // the model as one row per neuron (read by one-wave workgroups)
  bool have_model = false;
  GraphCache graph;
  std::vector<TuneEntry> tuned;
  double last_sift_ms = 0.0, last_mean_ms = 0.0;
  int last_threads = 0, last_grid = 0, last_ntask = 0;
  bool timing_pending = false;
  bool last_graph = false;
  int64_t graph_captures = 0, graph_launches = 0;
  int last_phase_grid = 0;
  // state of the spectra entry points (eemd_spectra.hip)
  void* spectra = nullptr;
  void (*spectra_free)(void*) = nullptr;
};

namespace {

// Kernel times of the last call from its events (a graph replay is timed
// as a whole and reported as the sift time).
hipError_t read_timing(eemd_ctx* c) {
  float a = 0.f, b = 0.f;
  hipError_t e = hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
  if (e == hipSuccess && !c->last_graph) e = hipEventElapsedTime(&b, c->ev[1], c->ev[2]);
  c->last_sift_ms = a;
  c->last_mean_ms = b;
  return e;
}

int fail(eemd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(eemd_ctx* c, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  (void)hipGetLastError();
  return fail(c, e == hipErrorOutOfMemory ? EEMD_ERR_OOM : EEMD_ERR_DEVICE, m);
}

#define HIPCHK(c, expr, what)                 \
  do {                                        \
    hipError_t _e = (expr);                   \
    if (_e != hipSuccess) return hip_fail(c, _e, what); \
  } while (0)

struct Plan {
  std::vector<SignalDesc> sig;
  std::vector<TaskDesc> task;
  std::vector<int32_t> order;
  std::vector<int64_t> start;  // prefix of lengths (mean kernel)
  int64_t total_n = 0;
  int64_t nmax = 0;
  int64_t acc_doubles = 0;
  int64_t units = 0;
  int threads = 0;
  int mpt = 0;
  int wave_windows = 0;
  int block = 0;  // block-Thomas block size the launch uses
  bool tuned = false;
  int grid = 0;
  int64_t npad = 0, cap = 0, rows = 0, cols = 0;
  int64_t stride = 0;
};

int auto_threads(int64_t nmax, bool fp64) {
  // Few members (the 64-thread rule in make_plan does not apply): a
  // workgroup per member.  1024-thread workgroups cap a lane at 128 VGPRs,
  // below what the fused sift pass needs; 512 keeps every FP32 value in
  // registers.  The FP64 pass holds twice the registers per sample and runs
  // fastest at 256 (measured on the 2^20-sample, 256-member workload: 425 ms
  // at 256, 525 ms at 512).
  if (nmax <= 8192 || fp64) return 256;
  return 512;
}

int validate(eemd_ctx* c, const int64_t* offsets, int32_t nsig, int32_t E, const double* sigmas,
             const eemd_params* p) {
  if (!c) return EEMD_ERR_BAD_ARG;
  if (!p) return fail(c, EEMD_ERR_BAD_ARG, "params is NULL");
  if (!offsets || nsig < 1) return fail(c, EEMD_ERR_BAD_ARG, "need at least one signal");
  if (E < 1) return fail(c, EEMD_ERR_BAD_ARG, "ensemble size E must be >= 1");
  if (!sigmas) return fail(c, EEMD_ERR_BAD_ARG, "sigmas is NULL");
  if (p->max_imf < 1 || p->max_imf > MAX_IMF_CAP)
    return fail(c, EEMD_ERR_BAD_ARG, "max_imf must be in [1, 64]");
  if (p->max_sifts < 0) return fail(c, EEMD_ERR_BAD_ARG, "max_sifts must be >= 0");
  if (p->end_mode < 0 || p->end_mode > 2) return fail(c, EEMD_ERR_BAD_ARG, "bad end_mode");
  if (p->end_mode == EEMD_END_LEARNED && !c->have_model)
    return fail(c, EEMD_ERR_BAD_ARG, "learned boundary needs a model (eemd_set_boundary_model)");
  if (p->halo < 0) return fail(c, EEMD_ERR_BAD_ARG, "halo must be >= 0");
  if (p->threads != 0 && p->threads != 64 && p->threads != 256 && p->threads != 512 &&
      p->threads != 1024)
    return fail(c, EEMD_ERR_BAD_ARG, "threads must be 0, 64, 256, 512 or 1024");
  if (p->members_per_task < 0) return fail(c, EEMD_ERR_BAD_ARG, "members_per_task must be >= 0");
  if (p->end_mode == EEMD_END_LEARNED && p->fp64 && p->threads == 1024)
    return fail(c, EEMD_ERR_BAD_ARG,
                "FP64 with learned ends needs threads 256 or 512 (LDS holds the model)");
  if (p->member0 < 0 || (int64_t)p->member0 + E > (int64_t(1) << 31))
    return fail(c, EEMD_ERR_BAD_ARG, "member0 must be >= 0 and member0 + E < 2^31");
  if (p->signal0 < 0 || (int64_t)p->signal0 + nsig > (int64_t(1) << 31))
    return fail(c, EEMD_ERR_BAD_ARG, "signal0 must be >= 0 and signal0 + nsig < 2^31");
  if (offsets[0] != 0) return fail(c, EEMD_ERR_BAD_ARG, "offsets[0] must be 0");
  for (int s = 0; s < nsig; ++s) {
    const int64_t n = offsets[s + 1] - offsets[s];
    if (n < 3) return fail(c, EEMD_ERR_BAD_ARG, "every signal needs at least 3 samples");
    if (n > (int64_t(1) << 30)) return fail(c, EEMD_ERR_BAD_ARG, "signal longer than 2^30");
    if (!p->fp64 && n > (int64_t(1) << 24))
      return fail(c, EEMD_ERR_BAD_ARG, "FP32 mode holds sample positions exactly only up to 2^24");
  }
  if ((int64_t)nsig * E > (int64_t(1) << 30)) return fail(c, EEMD_ERR_BAD_ARG, "too many members");
  return EEMD_OK;
}

// Shape of a batch as far as the launch configuration is concerned.
std::vector<int64_t> tune_key(const int64_t* offsets, int32_t nsig, int32_t E, const eemd_params* p) {
  int64_t nmax = 0;
  for (int s = 0; s < nsig; ++s) nmax = std::max<int64_t>(nmax, offsets[s + 1] - offsets[s]);
  return {p->fp64, p->end_mode, p->max_imf, p->max_sifts, p->block, p->halo, nsig, E,
          offsets[nsig], nmax};
}

int make_plan(eemd_ctx* c, const int64_t* offsets, int32_t nsig, int32_t E, const double* sigmas,
              const double* noise, const eemd_params* p, bool want_mimf, Plan& P) {
  const int tb = p->fp64 ? 8 : 4;
  P.nmax = 0;
  for (int s = 0; s < nsig; ++s) P.nmax = std::max<int64_t>(P.nmax, offsets[s + 1] - offsets[s]);
  int want_threads = p->threads, want_mpt = p->members_per_task, want_ww = p->wave_windows;
  P.tuned = false;
  P.block = p->block;
  if (want_threads == 0 && want_mpt == 0 && want_ww == 0) {
    const std::vector<int64_t> key = tune_key(offsets, nsig, E, p);
    for (const TuneEntry& t : c->tuned)
      if (t.key == key) {
        want_threads = t.threads;
        want_mpt = t.mpt;
        want_ww = t.wave_windows;
        // the tuned block size only where the caller allows it
        if (p->flags & EEMD_FLAG_TUNE_BLOCK) P.block = t.block;
        P.tuned = true;
        break;
      }
  }
  P.threads = want_threads;
  if (!P.threads) {
    // many members: one wave per member (64-thread workgroups, many per CU)
    // when the batch fills every wave slot of the chip at least twice over
    int b64 = 1;
    hipError_t oe = sift_occupancy(p->fp64, 64, p->end_mode, &b64);
    if (oe != hipSuccess) return hip_fail(c, oe, "occupancy query");
    if ((int64_t)nsig * E >= 2 * (int64_t)std::max(b64, 1) * c->ncu) P.threads = 64;
  }
  if (!P.threads) P.threads = auto_threads(P.nmax, p->fp64 != 0);
  const int nw = P.threads / WAVE;
  P.wave_windows = want_ww == 0 ? 8 * nw : (want_ww < 0 ? -1 : want_ww);

  int bpc = 1;
  hipError_t oe = sift_occupancy(p->fp64, P.threads, p->end_mode, &bpc);
  if (oe != hipSuccess) return hip_fail(c, oe, "occupancy query");
  if (bpc < 1) bpc = 1;
  const int64_t max_slots = (int64_t)bpc * c->ncu;

  // Members per task.  An explicit value applies to every signal.  The
  // automatic value is one count for the whole batch: the most members whose
  // task on the longest signal is at most one slot's fair share of the work
  // (W / slots sample-members, W = total sample-members).  Tasks of shorter
  // signals are proportionally smaller and come last in the longest-first
  // queue, where they fill the tail.  The accumulators (one max_imf x n_s
  // FP64 block per task) come to about slots x max_imf x n_max doubles.
  // A short member costs more than its sample count says (each sift's window
  // solve is a serial chain of up to B + 2S rows however short the signal),
  // so a rule that gave short signals few large tasks left those tasks to
  // run last and the launch idle for a quarter of its time on batches of
  // uneven lengths.
  int64_t work = 0;
  for (int s = 0; s < nsig; ++s) work += (offsets[s + 1] - offsets[s]) * (int64_t)E;
  const int mpt_auto = (int)std::max<int64_t>(
      1, std::min<int64_t>(work / (max_slots * std::max<int64_t>(P.nmax, 1)), E));
  auto mpt_of = [&](int64_t) -> int { return want_mpt > 0 ? std::min(want_mpt, E) : mpt_auto; };
  P.mpt = want_mpt > 0 ? std::min(want_mpt, E) : 0;

  P.sig.resize(nsig);
  P.start.resize(nsig);
  P.task.clear();
  int64_t noise_off = 0, avg_off = 0, mimf_off = 0, acc_off = 0, tot = 0;
  for (int s = 0; s < nsig; ++s) {
    const int64_t n = offsets[s + 1] - offsets[s];
    SignalDesc& d = P.sig[s];
    d.x_off = offsets[s];
    d.n = n;
    d.noise_off = noise ? noise_off : -1;
    d.avg_off = avg_off;
    d.mimf_off = want_mimf ? mimf_off : -1;
    d.sigma = sigmas[s];
    d.unit0 = s * E;
    d.task0 = (int32_t)P.task.size();
    d.nmember = E;
    const int mpt = mpt_of(n);
    for (int m0 = 0; m0 < E; m0 += mpt) {
      TaskDesc t{};
      t.signal = s;
      t.m0 = m0;
      t.m1 = std::min(E, m0 + mpt);
      t.acc_off = acc_off;
      acc_off += (int64_t)p->max_imf * n;
      P.task.push_back(t);
    }
    d.ntask = (int32_t)P.task.size() - d.task0;
    P.start[s] = tot;
    tot += n;
    noise_off += (int64_t)E * n;
    avg_off += (int64_t)p->max_imf * n;
    mimf_off += (int64_t)E * p->max_imf * n;
  }
  P.total_n = tot;
  P.acc_doubles = acc_off;
  P.units = (int64_t)nsig * E;
  P.order.resize(P.task.size());
  std::iota(P.order.begin(), P.order.end(), 0);
  std::stable_sort(P.order.begin(), P.order.end(), [&](int a, int b) {
    const auto& ta = P.task[a];
    const auto& tbk = P.task[b];
    const int64_t wa = P.sig[ta.signal].n * (ta.m1 - ta.m0);
    const int64_t wb = P.sig[tbk.signal].n * (tbk.m1 - tbk.m0);
    return wa > wb;
  });
  P.grid = (int)std::min<int64_t>((int64_t)P.task.size(), max_slots);

  // scratch geometry
  P.npad = ((P.nmax + CHUNK - 1) / CHUNK) * CHUNK + CHUNK;
  P.cap = seg_cap(seg_len(P.nmax, nw));
  if (P.block <= 0) {
    P.rows = P.npad / 2 + 2 * MAX_VIRT + 4;
    P.cols = 2;
  } else {
    P.rows = (int64_t)P.block + 2 * (int64_t)p->halo;
    // FP32: room for the wave-parallel window solve (3 x 64 x 64 floats per wave)
    if (!p->fp64) P.rows = std::max<int64_t>(P.rows, 3 * 64 * WAVE * 4 / (2 * 4 * WAVE));
    P.cols = P.threads;
  }
  P.stride = make_layout(tb, P.npad, nw, P.cap, P.rows, P.cols, p->tau > 0.0 ? 1 : 0).total;
  return EEMD_OK;
}

// Device buffers one call's kernels use besides the caller's inputs and
// outputs: the context's grow-only buffers, or carve-outs of the static
// allocation-free scratch.
struct DevPtrs {
  char* scratch = nullptr;  // per-workgroup slabs
  double* acc = nullptr;
  int32_t* acc_cnt = nullptr;
  int32_t* queue = nullptr;
  uint64_t* phase = nullptr;
  SignalDesc* sigd = nullptr;
  TaskDesc* taskd = nullptr;
  int32_t* order = nullptr;
  int64_t* sig_start = nullptr;
};

hipError_t ensure_desc(const Plan& P, DescBufs& D) {
  const size_t nt = P.task.size();
  hipError_t e;
  if ((e = D.sigd.ensure(sizeof(SignalDesc) * P.sig.size())) != hipSuccess) return e;
  if ((e = D.taskd.ensure(sizeof(TaskDesc) * nt)) != hipSuccess) return e;
  if ((e = D.order.ensure(sizeof(int32_t) * nt)) != hipSuccess) return e;
  return D.sig_start.ensure(sizeof(int64_t) * P.start.size());
}

void desc_ptrs(const DescBufs& D, DevPtrs& Q) {
  Q.sigd = (SignalDesc*)D.sigd.p;
  Q.taskd = (TaskDesc*)D.taskd.p;
  Q.order = (int32_t*)D.order.p;
  Q.sig_start = (int64_t*)D.sig_start.p;
}

hipError_t upload_desc(const Plan& P, const DevPtrs& D, hipStream_t st) {
  const size_t nt = P.task.size();
  hipError_t e;
  if ((e = hipMemcpyAsync(D.sigd, P.sig.data(), sizeof(SignalDesc) * P.sig.size(),
                          hipMemcpyHostToDevice, st)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(D.taskd, P.task.data(), sizeof(TaskDesc) * nt,
                          hipMemcpyHostToDevice, st)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(D.order, P.order.data(), sizeof(int32_t) * nt,
                          hipMemcpyHostToDevice, st)) != hipSuccess) return e;
  return hipMemcpyAsync(D.sig_start, P.start.data(), sizeof(int64_t) * P.start.size(),
                        hipMemcpyHostToDevice, st);
}

// The device work of one call: clears (fill kernels, never memsets, so the
// sequence can be captured), the persistent sift kernel and the mean kernel.
// With `events`, ev[0..2] bracket the sift and mean kernels.
hipError_t enqueue_work(eemd_ctx* c, const Plan& P, SiftArgs A, const eemd_params* p,
                        const DevPtrs& D, const double* x, double* avg, double* resid,
                        int32_t* n_modes, hipStream_t st, bool events) {
  hipError_t e;
  A.sig = D.sigd;
  A.task = D.taskd;
  A.order = D.order;
  if ((e = launch_fill_i32(D.queue, 1, 0, st)) != hipSuccess) return e;
  if ((e = launch_fill_f64(avg, (int64_t)p->max_imf * P.total_n, 0.0, st)) != hipSuccess) return e;
  const int tb = p->fp64 ? 8 : 4;
  if (A.member_imfs) {
    int64_t bytes = 0;
    for (const auto& s : P.sig) bytes += (int64_t)s.nmember * p->max_imf * s.n * tb;
    if ((e = launch_fill_bytes(A.member_imfs, bytes, st)) != hipSuccess) return e;
  }
  if (A.trace) {
    const int64_t words = P.units * p->max_imf * std::max(p->max_sifts, 1) * 4;
    if ((e = launch_fill_bytes(A.trace, words * 8, st)) != hipSuccess) return e;
  }
  if (events && (e = hipEventRecord(c->ev[0], st)) != hipSuccess) return e;
  if ((e = launch_sift(A, p->fp64, P.threads, P.grid, st)) != hipSuccess) return e;
  if (events && (e = hipEventRecord(c->ev[1], st)) != hipSuccess) return e;
  if ((e = launch_mean(D.sigd, (int)P.sig.size(), D.sig_start, D.taskd, D.acc, D.acc_cnt,
                       p->max_imf, x, avg, resid, n_modes, P.total_n,
                       (p->flags & EEMD_FLAG_LEVEL_SUMS) ? 1 : 0, st)) != hipSuccess)
    return e;
  if (events && (e = hipEventRecord(c->ev[2], st)) != hipSuccess) return e;
  return hipSuccess;
}

template <typename V>
void key_put(std::vector<uint8_t>& k, const V& v) {
  const uint8_t* b = reinterpret_cast<const uint8_t*>(&v);
  k.insert(k.end(), b, b + sizeof(V));
}

template <typename V>
void key_put_vec(std::vector<uint8_t>& k, const std::vector<V>& v) {
  key_put(k, v.size());
  const uint8_t* b = reinterpret_cast<const uint8_t*>(v.data());
  k.insert(k.end(), b, b + sizeof(V) * v.size());
}

// Kernel arguments of one call (everything but the descriptor pointers,
// which enqueue_work fills in).
SiftArgs sift_args(const Plan& P, const DevPtrs& D, const double* x, const double* noise,
                   const eemd_params* p, int32_t* nmode, int32_t* nsift, int32_t* per_mode,
                   void* mimf, uint64_t* trace, const void* mlp, const void* mlp_img) {
  SiftArgs A;
  std::memset(&A, 0, sizeof(A));  // padding included: A is part of the graph key
  A.x = x;
  A.noise = noise;
  A.ntask = (int32_t)P.task.size();
  A.nsig = (int32_t)P.sig.size();
  A.queue = D.queue;
  A.seed = p->seed;
  A.max_imf = p->max_imf;
  A.max_sifts = p->max_sifts;
  A.min_extrema = p->min_extrema;
  A.end_mode = p->end_mode;
  A.tau = p->tau;
  A.B = P.block;
  A.S = p->halo;
  A.acc = D.acc;
  A.acc_cnt = D.acc_cnt;
  A.member_imfs = mimf;
  A.nmode = nmode;
  A.nsift = nsift;
  A.per_mode = per_mode;
  A.trace = trace;
  A.scratch = D.scratch;
  A.scratch_stride = P.stride;
  A.npad = P.npad;
  A.cap = P.cap;
  A.rows = P.rows;
  A.cols = P.cols;
  A.mlp = (p->end_mode == EEMD_END_LEARNED) ? mlp : nullptr;
  A.mlp_img = (p->end_mode == EEMD_END_LEARNED) ? mlp_img : nullptr;
  A.nslots = P.grid;
  A.phase = (p->flags & EEMD_FLAG_PHASES) ? D.phase : nullptr;
  A.wave_windows = P.wave_windows;
  A.member0 = p->member0;
  A.signal0 = p->signal0;
  return A;
}

int run_plan(eemd_ctx* c, const Plan& P, const double* x, const double* noise,
             const eemd_params* p, double* avg, double* resid, int32_t* n_modes, int32_t* nmode,
             int32_t* nsift, int32_t* per_mode, void* mimf, uint64_t* trace, hipStream_t st,
             bool sync_timing) {
  const size_t nt = P.task.size();
  HIPCHK(c, c->scratch.ensure((size_t)P.stride * P.grid), "scratch alloc");
  HIPCHK(c, c->acc.ensure(sizeof(double) * (size_t)std::max<int64_t>(P.acc_doubles, 1)), "acc alloc");
  HIPCHK(c, c->acc_cnt.ensure(sizeof(int32_t) * nt * p->max_imf), "acc_cnt alloc");
  HIPCHK(c, c->queue.ensure(256), "queue alloc");
  if (p->flags & EEMD_FLAG_PHASES)
    HIPCHK(c, c->phase.ensure(sizeof(uint64_t) * 4 * P.grid), "phase alloc");
  DevPtrs D;
  D.scratch = (char*)c->scratch.p;
  D.acc = (double*)c->acc.p;
  D.acc_cnt = (int32_t*)c->acc_cnt.p;
  D.queue = (int32_t*)c->queue.p;
  D.phase = (uint64_t*)c->phase.p;
  SiftArgs A = sift_args(P, D, x, noise, p, nmode, nsift, per_mode, mimf, trace, c->mlpw.p,
                         c->mlpi.p);
  c->last_phase_grid = (p->flags & EEMD_FLAG_PHASES) ? P.grid : 0;
  // order this call after the previous one, whatever stream it used
  if (c->done_recorded) HIPCHK(c, hipStreamWaitEvent(st, c->done, 0), "stream wait");

  if (p->flags & EEMD_FLAG_GRAPH) {
    // everything the captured sequence depends on: the plan, the parameters
    // and every device address it was captured with
    std::vector<uint8_t> key;
    key_put_vec(key, P.sig);
    key_put_vec(key, P.task);
    key_put_vec(key, P.order);
    key_put_vec(key, P.start);
    eemd_params pk = *p;
    key_put(key, pk);
    key_put(key, P.threads);
    key_put(key, P.grid);
    key_put(key, P.stride);
    key_put(key, P.wave_windows);
    key_put(key, P.block);
    key_put(key, A);
    // the mean kernel's outputs are launch arguments outside SiftArgs
    key_put(key, avg);
    key_put(key, resid);
    key_put(key, n_modes);
    GraphCache& G = c->graph;
    if (!G.exec || G.key != key) {
      G.reset();
      hipStream_t cs = c->stream;
      if (st != cs) HIPCHK(c, hipStreamSynchronize(st), "stream sync before capture");
      HIPCHK(c, ensure_desc(P, G.desc), "descriptor alloc");
      DevPtrs DG = D;
      desc_ptrs(G.desc, DG);
      HIPCHK(c, upload_desc(P, DG, cs), "descriptor upload");
      HIPCHK(c, hipStreamSynchronize(cs), "descriptor upload");
      HIPCHK(c, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
      hipError_t e = enqueue_work(c, P, A, p, DG, x, avg, resid, n_modes, cs, false);
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(cs, &g);
      if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(c, e, "graph capture");
      }
      if (e2 != hipSuccess) return hip_fail(c, e2, "end capture");
      e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        G.exec = nullptr;
        return hip_fail(c, e, "graph instantiate");
      }
      G.key = key;
      ++c->graph_captures;
    }
    HIPCHK(c, hipEventRecord(c->ev[0], st), "event");
    HIPCHK(c, hipGraphLaunch(G.exec, st), "graph launch");
    HIPCHK(c, hipEventRecord(c->ev[1], st), "event");
    HIPCHK(c, hipEventRecord(c->ev[2], st), "event");
    c->last_graph = true;
    ++c->graph_launches;
  } else {
    HIPCHK(c, ensure_desc(P, c->desc), "descriptor alloc");
    desc_ptrs(c->desc, D);
    HIPCHK(c, upload_desc(P, D, st), "descriptor upload");
    HIPCHK(c, enqueue_work(c, P, A, p, D, x, avg, resid, n_modes, st, true), "kernel launch");
    c->last_graph = false;
  }
  HIPCHK(c, hipEventRecord(c->done, st), "event");
  c->done_recorded = true;
  c->last_threads = P.threads;
  c->last_grid = P.grid;
  c->last_ntask = (int)nt;
  c->timing_pending = !sync_timing;
  if (sync_timing) {
    HIPCHK(c, hipStreamSynchronize(st), "kernel execution");
    HIPCHK(c, read_timing(c), "event timing");
  }
  return EEMD_OK;
}

// The allocation-free form of run_plan: every device buffer of the call
// besides the caller's comes from this device's `scratch` array (its address
// from noalloc_lookup).  The plan's grid shrinks to the slabs that fit; when
// the accumulators of the automatic task split do not fit, every signal
// becomes one task.  No graph capture, no buffer growth, no hipMalloc.
int run_plan_noalloc(eemd_ctx* c, const int64_t* offsets, int32_t nsig, int32_t E,
                     const double* sigmas, const double* x, const double* noise,
                     const eemd_params* p, double* avg, double* resid, int32_t* n_modes,
                     int32_t* nmode, int32_t* nsift, int32_t* per_mode, void* mimf,
                     uint64_t* trace, hipStream_t st, bool sync_timing) {
  const int dev = c->device;
  if (dev < 0 || dev >= 64 || !g_noalloc_done[dev])
    return fail(c, EEMD_ERR_DEVICE, "allocation-free scratch: device not set up");
  if (!g_noalloc_ptr[dev])
    return hip_fail(c, g_noalloc_err[dev], "allocation-free scratch: hipGetSymbolAddress");
  char* const base = static_cast<char*>(g_noalloc_ptr[dev]);
  const int64_t cap = (int64_t)EEMD_NOALLOC_SCRATCH_BYTES;
  Plan P;
  int rc = make_plan(c, offsets, nsig, E, sigmas, noise, p, mimf != nullptr, P);
  if (rc) return rc;
  // bytes of everything but the slabs, for a plan
  auto fixed_bytes = [&](const Plan& Q) -> int64_t {
    const int64_t nt = (int64_t)Q.task.size();
    int64_t b = 256;                                                   // queue
    b += align256((int64_t)sizeof(SignalDesc) * nsig);
    b += align256((int64_t)sizeof(TaskDesc) * nt);
    b += align256(4 * nt);                                             // order
    b += align256(8 * (int64_t)nsig);                                  // signal starts
    b += align256(4 * nt * p->max_imf);                                // level counts
    if (!per_mode) b += align256(4 * Q.units * p->max_imf);
    if (p->flags & EEMD_FLAG_PHASES) b += align256(32 * (int64_t)Q.grid);
    b += align256(8 * std::max<int64_t>(Q.acc_doubles, 1));            // accumulators
    return b;
  };
  if (fixed_bytes(P) + P.stride > cap && p->members_per_task == 0) {
    eemd_params q = *p;
    q.members_per_task = E;  // one task (one accumulator block) per signal
    rc = make_plan(c, offsets, nsig, E, sigmas, noise, &q, mimf != nullptr, P);
    if (rc) return rc;
  }
  const int64_t fixed = fixed_bytes(P);
  if (fixed + P.stride > cap) {
    char msg[256];
    std::snprintf(msg, sizeof msg,
                  "batch too large for the allocation-free scratch: needs %lld bytes, "
                  "EEMD_NOALLOC_SCRATCH_BYTES is %lld",
                  (long long)(fixed + P.stride), (long long)cap);
    return fail(c, EEMD_ERR_BAD_ARG, msg);
  }
  P.grid = (int)std::min<int64_t>(P.grid, (cap - fixed) / P.stride);
  const size_t nt = P.task.size();
  // carve the scratch
  DevPtrs D;
  int64_t o = 0;
  auto take = [&](int64_t bytes) -> char* {
    char* q = base + o;
    o += align256(bytes);
    return q;
  };
  D.queue = (int32_t*)take(256);
  D.sigd = (SignalDesc*)take((int64_t)sizeof(SignalDesc) * nsig);
  D.taskd = (TaskDesc*)take((int64_t)sizeof(TaskDesc) * nt);
  D.order = (int32_t*)take(4 * (int64_t)nt);
  D.sig_start = (int64_t*)take(8 * (int64_t)nsig);
  D.acc_cnt = (int32_t*)take(4 * (int64_t)nt * p->max_imf);
  int32_t* pm = per_mode ? per_mode : (int32_t*)take(4 * P.units * p->max_imf);
  if (p->flags & EEMD_FLAG_PHASES) D.phase = (uint64_t*)take(32 * (int64_t)P.grid);
  D.acc = (double*)take(8 * std::max<int64_t>(P.acc_doubles, 1));
  D.scratch = take(P.stride * P.grid);
  if (o > cap) return fail(c, EEMD_ERR_BAD_ARG, "allocation-free scratch overflow");
  SiftArgs A = sift_args(P, D, x, noise, p, nmode, nsift, pm, mimf, trace, c->mlpw.p,
                         c->mlpi.p);
  c->last_phase_grid = 0;  // eemd_last_phases reads the context's phase buffer
  if (c->done_recorded) HIPCHK(c, hipStreamWaitEvent(st, c->done, 0), "stream wait");
  HIPCHK(c, upload_desc(P, D, st), "descriptor upload");
  HIPCHK(c, enqueue_work(c, P, A, p, D, x, avg, resid, n_modes, st, true), "kernel launch");
  c->last_graph = false;
  HIPCHK(c, hipEventRecord(c->done, st), "event");
  c->done_recorded = true;
  c->last_threads = P.threads;
  c->last_grid = P.grid;
  c->last_ntask = (int)nt;
  c->timing_pending = !sync_timing;
  if (sync_timing) {
    HIPCHK(c, hipStreamSynchronize(st), "kernel execution");
    HIPCHK(c, read_timing(c), "event timing");
  }
  return EEMD_OK;
}

}  // namespace

extern "C" {

void eemd_params_default(eemd_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->fp64 = 1;
  p->end_mode = EEMD_END_MIRROR;
  p->max_imf = 10;
  p->max_sifts = 10;
  p->min_extrema = 4;
  p->block = 1024;
  p->halo = 30;
  p->threads = 0;
  p->members_per_task = 0;
  p->flags = 0;
  p->tau = 0.0;
  p->seed = 0;
}

int eemd_ctx_create(int device, eemd_ctx** out) {
  if (!out) return EEMD_ERR_BAD_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    (void)hipGetLastError();
    return EEMD_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev) return EEMD_ERR_BAD_ARG;
  eemd_ctx* c = new eemd_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipStreamCreate(&c->stream) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return EEMD_ERR_DEVICE;
  }
  for (hipEvent_t* e : {&c->ev[0], &c->ev[1], &c->ev[2], &c->done}) {
    if (hipEventCreate(e) != hipSuccess) {
      (void)hipGetLastError();
      delete c;
      return EEMD_ERR_DEVICE;
    }
  }
  // the allocation-free scratch of this device (device current)
  noalloc_lookup(device);
  *out = c;
  return EEMD_OK;
}

int eemd_ctx_destroy(eemd_ctx* c) {
  if (!c) return EEMD_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->graph.reset();
  c->graph.desc.release();
  if (c->spectra && c->spectra_free) c->spectra_free(c->spectra);
  c->spectra = nullptr;
  c->desc.release();
  for (DevBuf* b : {&c->scratch, &c->acc, &c->acc_cnt, &c->queue, &c->phase, &c->hx, &c->hnoise,
                    &c->havg, &c->hres, &c->hnm, &c->hnmode, &c->hnsift, &c->hpm, &c->hmi, &c->htr,
                    &c->tx, &c->tavg, &c->tres, &c->tnm, &c->tnmode, &c->tnsift, &c->tpm,
                    &c->mlpw, &c->mlpi})
    b->release();
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->done) (void)hipEventDestroy(c->done);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return EEMD_OK;
}

const char* eemd_last_error(const eemd_ctx* c) { return c ? c->err.c_str() : "null context"; }

int eemd_run_batch_dev(eemd_ctx* c, const double* x, const int64_t* offsets, int32_t nsig,
                       int32_t E, const double* sigmas, const double* noise,
                       const eemd_params* p, double* avg, double* resid, int32_t* n_modes,
                       int32_t* nmode, int32_t* nsift, int32_t* per_mode, void* mimf,
                       uint64_t* trace, void* stream) {
  int rc = validate(c, offsets, nsig, E, sigmas, p);
  if (rc) return rc;
  if (!x || !avg || !resid || !n_modes || !nmode || !nsift)
    return fail(c, EEMD_ERR_BAD_ARG, "required device buffer is NULL");
  if (mimf && !(p->flags & EEMD_FLAG_MEMBER_IMFS))
    return fail(c, EEMD_ERR_BAD_ARG, "member_imfs given without EEMD_FLAG_MEMBER_IMFS");
  if (trace && !(p->flags & EEMD_FLAG_TRACE))
    return fail(c, EEMD_ERR_BAD_ARG, "trace given without EEMD_FLAG_TRACE");
  HIPCHK(c, hipSetDevice(c->device), "set device");
  Plan P;
  rc = make_plan(c, offsets, nsig, E, sigmas, noise, p, mimf != nullptr, P);
  if (rc) return rc;
  // per_mode is required by the kernel; use a context buffer when absent
  int32_t* pm = per_mode;
  if (!pm) {
    HIPCHK(c, c->hpm.ensure(sizeof(int32_t) * P.units * p->max_imf), "per-mode alloc");
    pm = (int32_t*)c->hpm.p;
  }
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  return run_plan(c, P, x, noise, p, avg, resid, n_modes, nmode, nsift, pm, mimf, trace, st,
                  stream == nullptr);
}

int eemd_run_batch_dev_noalloc(eemd_ctx* c, const double* x, const int64_t* offsets,
                               int32_t nsig, int32_t E, const double* sigmas, const double* noise,
                               const eemd_params* p, double* avg, double* resid, int32_t* n_modes,
                               int32_t* nmode, int32_t* nsift, int32_t* per_mode, void* mimf,
                               uint64_t* trace, void* stream) {
  int rc = validate(c, offsets, nsig, E, sigmas, p);
  if (rc) return rc;
  if (!x || !avg || !resid || !n_modes || !nmode || !nsift)
    return fail(c, EEMD_ERR_BAD_ARG, "required device buffer is NULL");
  if (mimf && !(p->flags & EEMD_FLAG_MEMBER_IMFS))
    return fail(c, EEMD_ERR_BAD_ARG, "member_imfs given without EEMD_FLAG_MEMBER_IMFS");
  if (trace && !(p->flags & EEMD_FLAG_TRACE))
    return fail(c, EEMD_ERR_BAD_ARG, "trace given without EEMD_FLAG_TRACE");
  if (p->flags & EEMD_FLAG_GRAPH)
    return fail(c, EEMD_ERR_BAD_ARG, "EEMD_FLAG_GRAPH is not available without allocation");
  HIPCHK(c, hipSetDevice(c->device), "set device");
  hipStream_t st = stream ? (hipStream_t)stream : c->stream;
  return run_plan_noalloc(c, offsets, nsig, E, sigmas, x, noise, p, avg, resid, n_modes, nmode,
                          nsift, per_mode, mimf, trace, st, stream == nullptr);
}

int eemd_run_batch(eemd_ctx* c, const double* x, const int64_t* offsets, int32_t nsig, int32_t E,
                   const double* sigmas, const double* noise, const eemd_params* p, double* avg,
                   double* resid, int32_t* n_modes, int32_t* nmode, int32_t* nsift,
                   int32_t* per_mode, void* mimf, uint64_t* trace) {
  int rc = validate(c, offsets, nsig, E, sigmas, p);
  if (rc) return rc;
  if (!x || !avg || !resid || !n_modes || !nmode || !nsift)
    return fail(c, EEMD_ERR_BAD_ARG, "required output buffer is NULL");
  if (mimf && !(p->flags & EEMD_FLAG_MEMBER_IMFS))
    return fail(c, EEMD_ERR_BAD_ARG, "member_imfs given without EEMD_FLAG_MEMBER_IMFS");
  if (trace && !(p->flags & EEMD_FLAG_TRACE))
    return fail(c, EEMD_ERR_BAD_ARG, "trace given without EEMD_FLAG_TRACE");
  HIPCHK(c, hipSetDevice(c->device), "set device");
  Plan P;
  rc = make_plan(c, offsets, nsig, E, sigmas, noise, p, mimf != nullptr, P);
  if (rc) return rc;
  const int tb = p->fp64 ? 8 : 4;
  const int64_t ntot = P.total_n;
  const int64_t navg = (int64_t)p->max_imf * ntot;
  int64_t nmi = 0;
  for (const auto& s : P.sig) nmi += (int64_t)E * p->max_imf * s.n;
  const int64_t ntr = P.units * p->max_imf * std::max(p->max_sifts, 1) * 4;
  hipStream_t st = c->stream;
  HIPCHK(c, c->hx.ensure(sizeof(double) * ntot), "input alloc");
  if (noise) HIPCHK(c, c->hnoise.ensure(sizeof(double) * ntot * E), "noise alloc");
  HIPCHK(c, c->havg.ensure(sizeof(double) * navg), "output alloc");
  HIPCHK(c, c->hres.ensure(sizeof(double) * ntot), "output alloc");
  HIPCHK(c, c->hnm.ensure(sizeof(int32_t) * nsig), "output alloc");
  HIPCHK(c, c->hnmode.ensure(sizeof(int32_t) * P.units), "output alloc");
  HIPCHK(c, c->hnsift.ensure(sizeof(int32_t) * P.units), "output alloc");
  HIPCHK(c, c->hpm.ensure(sizeof(int32_t) * P.units * p->max_imf), "output alloc");
  if (mimf) HIPCHK(c, c->hmi.ensure((size_t)tb * nmi), "member imf alloc");
  if (trace) HIPCHK(c, c->htr.ensure(sizeof(uint64_t) * ntr), "trace alloc");
  HIPCHK(c, hipMemcpyAsync(c->hx.p, x, sizeof(double) * ntot, hipMemcpyHostToDevice, st), "H2D x");
  if (noise)
    HIPCHK(c, hipMemcpyAsync(c->hnoise.p, noise, sizeof(double) * ntot * E, hipMemcpyHostToDevice, st),
           "H2D noise");
  rc = run_plan(c, P, (const double*)c->hx.p, noise ? (const double*)c->hnoise.p : nullptr, p,
                (double*)c->havg.p, (double*)c->hres.p, (int32_t*)c->hnm.p, (int32_t*)c->hnmode.p,
                (int32_t*)c->hnsift.p, (int32_t*)c->hpm.p, mimf ? c->hmi.p : nullptr,
                trace ? (uint64_t*)c->htr.p : nullptr, st, true);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(avg, c->havg.p, sizeof(double) * navg, hipMemcpyDeviceToHost, st), "D2H");
  HIPCHK(c, hipMemcpyAsync(resid, c->hres.p, sizeof(double) * ntot, hipMemcpyDeviceToHost, st), "D2H");
  HIPCHK(c, hipMemcpyAsync(n_modes, c->hnm.p, sizeof(int32_t) * nsig, hipMemcpyDeviceToHost, st), "D2H");
  HIPCHK(c, hipMemcpyAsync(nmode, c->hnmode.p, sizeof(int32_t) * P.units, hipMemcpyDeviceToHost, st), "D2H");
  HIPCHK(c, hipMemcpyAsync(nsift, c->hnsift.p, sizeof(int32_t) * P.units, hipMemcpyDeviceToHost, st), "D2H");
  if (per_mode)
    HIPCHK(c, hipMemcpyAsync(per_mode, c->hpm.p, sizeof(int32_t) * P.units * p->max_imf,
                             hipMemcpyDeviceToHost, st), "D2H");
  if (mimf)
    HIPCHK(c, hipMemcpyAsync(mimf, c->hmi.p, (size_t)tb * nmi, hipMemcpyDeviceToHost, st), "D2H");
  if (trace)
    HIPCHK(c, hipMemcpyAsync(trace, c->htr.p, sizeof(uint64_t) * ntr, hipMemcpyDeviceToHost, st), "D2H");
  HIPCHK(c, hipStreamSynchronize(st), "D2H sync");
  return EEMD_OK;
}

int eemd_autotune(eemd_ctx* c, const double* x, const int64_t* offsets, int32_t nsig, int32_t E,
                  const double* sigmas, const double* noise, const eemd_params* prm, int32_t reps,
                  eemd_params* best, double* best_ms) {
  int rc = validate(c, offsets, nsig, E, sigmas, prm);
  if (rc) return rc;
  if (!x || !best) return fail(c, EEMD_ERR_BAD_ARG, "x and best are required");
  if (reps < 1) reps = 1;
  HIPCHK(c, hipSetDevice(c->device), "set device");
  const int64_t ntot = offsets[nsig];
  const int64_t units = (int64_t)nsig * E;
  HIPCHK(c, c->tavg.ensure(sizeof(double) * prm->max_imf * ntot), "tune alloc");
  HIPCHK(c, c->tres.ensure(sizeof(double) * ntot), "tune alloc");
  HIPCHK(c, c->tnm.ensure(sizeof(int32_t) * nsig), "tune alloc");
  HIPCHK(c, c->tnmode.ensure(sizeof(int32_t) * units), "tune alloc");
  HIPCHK(c, c->tnsift.ensure(sizeof(int32_t) * units), "tune alloc");
  HIPCHK(c, c->tpm.ensure(sizeof(int32_t) * units * prm->max_imf), "tune alloc");

  // kernel time of one configuration (explicit threads / mpt / wave_windows / block)
  auto timed = [&](int threads, int mpt, int ww, int block, double& ms) -> int {
    eemd_params q = *prm;
    q.threads = threads;
    q.members_per_task = mpt;
    q.wave_windows = ww;
    q.block = block;
    q.flags = 0;
    Plan P;
    int r = make_plan(c, offsets, nsig, E, sigmas, noise, &q, false, P);
    if (r) return r;
    ms = 1e300;
    for (int k = 0; k < reps; ++k) {
      r = run_plan(c, P, x, noise, &q, (double*)c->tavg.p, (double*)c->tres.p, (int32_t*)c->tnm.p,
                   (int32_t*)c->tnmode.p, (int32_t*)c->tnsift.p, (int32_t*)c->tpm.p, nullptr,
                   nullptr, c->stream, true);
      if (r) return r;
      ms = std::min(ms, c->last_sift_ms + c->last_mean_ms);
    }
    return EEMD_OK;
  };
  // accumulator bytes of a configuration (tasks x max_imf x n x 8)
  auto acc_bytes = [&](int threads, int mpt) -> int64_t {
    eemd_params q = *prm;
    q.threads = threads;
    q.members_per_task = mpt;
    q.wave_windows = -1;
    Plan P;
    if (make_plan(c, offsets, nsig, E, sigmas, noise, &q, false, P)) return INT64_MAX;
    return P.acc_doubles * 8;
  };

  // 1. workgroup size (automatic members per task, default wave threshold)
  int bt = 0, bm = 0, bw = 0, bb = prm->block;
  double bms = 1e300;
  for (int t : {64, 256, 512, 1024}) {
    // candidates validate() rejects are skipped (FP64 learned ends: the
    // model and the 1024-thread workgroup's LDS do not fit together)
    if (prm->fp64 && prm->end_mode == EEMD_END_LEARNED && t == 1024) continue;
    const int w = prm->fp64 ? -1 : 8 * (t / WAVE);
    double ms;
    rc = timed(t, 0, w, bb, ms);
    if (rc) return rc;
    if (ms < bms) { bms = ms; bt = t; bm = 0; bw = w; }
  }
  // 2. block size, where the caller allows it (results move by the seam
  // truncation only); exact Thomas (block <= 0) stays exact
  if ((prm->flags & EEMD_FLAG_TUNE_BLOCK) && prm->block > 0) {
    for (int B : {128, 256, 512, 1024, 2048}) {
      if (B == prm->block) continue;
      double ms;
      rc = timed(bt, 0, bw, B, ms);
      if (rc) return rc;
      if (ms < bms) { bms = ms; bb = B; }
    }
  }
  // 3. members per task: uniform choices, skipping those that leave slots
  // idle or need far more accumulator memory than the automatic rule
  {
    int bpc = 1;
    HIPCHK(c, sift_occupancy(prm->fp64, bt, prm->end_mode, &bpc), "occupancy query");
    const int sl = std::max(1, bpc) * c->ncu;
    const int64_t auto_acc = acc_bytes(bt, 0);
    std::vector<int> cand;
    for (int m : {1, 2, 4, 8, 16, E})
      if (m <= E && std::find(cand.begin(), cand.end(), m) == cand.end()) cand.push_back(m);
    for (int m : cand) {
      const int64_t ntask = (int64_t)nsig * ((E + m - 1) / m);
      if (ntask < sl) continue;
      const int64_t ab = acc_bytes(bt, m);
      if (ab > 2 * auto_acc && ab > (int64_t(4) << 30)) continue;
      double ms;
      rc = timed(bt, m, bw, bb, ms);
      if (rc) return rc;
      if (ms < bms) { bms = ms; bm = m; }
    }
  }
  // 4. FP32 wave-solve threshold
  if (!prm->fp64) {
    const int nw = bt / WAVE;
    for (int w : {-1, 2 * nw, 32 * nw}) {
      if (w == bw) continue;
      double ms;
      rc = timed(bt, bm, w, bb, ms);
      if (rc) return rc;
      if (ms < bms) { bms = ms; bw = w; }
    }
  }
  const std::vector<int64_t> key = tune_key(offsets, nsig, E, prm);
  bool found = false;
  for (TuneEntry& t : c->tuned)
    if (t.key == key) {
      t.threads = bt;
      t.mpt = bm;
      t.wave_windows = bw;
      t.block = bb;
      found = true;
    }
  if (!found) c->tuned.push_back(TuneEntry{key, bt, bm, bw, bb});
  *best = *prm;
  best->threads = bt;
  best->members_per_task = bm;
  best->wave_windows = bw;
  best->block = bb;
  if (best_ms) *best_ms = bms;
  return EEMD_OK;
}

int eemd_noise(eemd_ctx* c, uint64_t seed, int32_t signal, int32_t member0, int32_t E, int64_t n,
               double* out) {
  if (!c) return EEMD_ERR_BAD_ARG;
  if (!out || E < 1 || n < 1 || signal < 0 || member0 < 0)
    return fail(c, EEMD_ERR_BAD_ARG, "bad noise request");
  HIPCHK(c, hipSetDevice(c->device), "set device");
  DevBuf tmp;
  HIPCHK(c, tmp.ensure(sizeof(double) * (size_t)n * E), "noise alloc");
  hipError_t e = launch_noise(seed, (uint32_t)signal, (uint32_t)member0, E, n, (double*)tmp.p, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp.p, sizeof(double) * (size_t)n * E, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  tmp.release();
  if (e != hipSuccess) return hip_fail(c, e, "noise");
  return EEMD_OK;
}

int eemd_last_timing(eemd_ctx* c, double* sift_ms, double* mean_ms, int32_t* threads,
                     int32_t* grid, int32_t* ntasks) {
  if (!c) return EEMD_ERR_BAD_ARG;
  if (c->timing_pending) {
    // the last call was enqueued on a caller stream: wait for its events
    HIPCHK(c, hipEventSynchronize(c->ev[2]), "event wait");
    HIPCHK(c, read_timing(c), "event timing");
    c->timing_pending = false;
  }
  if (sift_ms) *sift_ms = c->last_sift_ms;
  if (mean_ms) *mean_ms = c->last_mean_ms;
  if (threads) *threads = c->last_threads;
  if (grid) *grid = c->last_grid;
  if (ntasks) *ntasks = c->last_ntask;
  return EEMD_OK;
}

int eemd_set_boundary_model(eemd_ctx* c, const float* params) {
  if (!c) return EEMD_ERR_BAD_ARG;
  if (!params) return fail(c, EEMD_ERR_BAD_ARG, "params is NULL");
  for (int i = 0; i < EEMD_MLP_NPARAM; ++i)
    if (!std::isfinite(params[i])) return fail(c, EEMD_ERR_BAD_ARG, "non-finite model parameter");
  std::vector<uint16_t> w(EEMD_MLP_NPARAM);
  for (int i = 0; i < EEMD_MLP_NPARAM; ++i) w[i] = f32_to_bf16(params[i]);
  // one row per neuron, W^T[out][in + pad] per layer, then the biases
  std::vector<uint16_t> img(MLP_LDS_N, 0);
  {
    const int gw[4] = {MLP_W1, MLP_W2, MLP_W3, MLP_W4};
    const int gb[4] = {MLP_B1, MLP_B2, MLP_B3, MLP_B4};
    const int lw[4] = {MLP_LDS_W1, MLP_LDS_W2, MLP_LDS_W3, MLP_LDS_W4};
    const int lb[4] = {MLP_LDS_B1, MLP_LDS_B2, MLP_LDS_B3, MLP_LDS_B4};
    for (int l = 0; l < 4; ++l) {
      const int nin = l == 0 ? MLP_IN : MLP_HID, nout = l == 3 ? MLP_OUT : MLP_HID;
      for (int i = 0; i < nin; ++i)
        for (int j = 0; j < nout; ++j) img[lw[l] + j * (nin + MLP_LDS_PAD) + i] = w[gw[l] + i * nout + j];
      for (int j = 0; j < nout; ++j) img[lb[l] + j] = w[gb[l] + j];
    }
  }
  HIPCHK(c, hipSetDevice(c->device), "set device");
  HIPCHK(c, c->mlpw.ensure(sizeof(uint16_t) * EEMD_MLP_NPARAM), "model alloc");
  HIPCHK(c, c->mlpi.ensure(sizeof(uint16_t) * MLP_LDS_N), "model alloc");
  HIPCHK(c, hipStreamSynchronize(c->stream), "model upload");
  HIPCHK(c, hipMemcpy(c->mlpw.p, w.data(), sizeof(uint16_t) * EEMD_MLP_NPARAM,
                      hipMemcpyHostToDevice), "model upload");
  HIPCHK(c, hipMemcpy(c->mlpi.p, img.data(), sizeof(uint16_t) * MLP_LDS_N,
                      hipMemcpyHostToDevice), "model upload");
  c->have_model = true;
  return EEMD_OK;
}

int eemd_mem_info(eemd_ctx* c, uint64_t* free_bytes, uint64_t* total_bytes) {
  if (!c) return EEMD_ERR_BAD_ARG;
  HIPCHK(c, hipSetDevice(c->device), "set device");
  size_t f = 0, t = 0;
  HIPCHK(c, hipMemGetInfo(&f, &t), "memory info");
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return EEMD_OK;
}

int eemd_graph_stats(eemd_ctx* c, int64_t* captures, int64_t* launches) {
  if (!c) return EEMD_ERR_BAD_ARG;
  if (captures) *captures = c->graph_captures;
  if (launches) *launches = c->graph_launches;
  return EEMD_OK;
}

int eemd_last_phases(eemd_ctx* c, double* ms4) {
  if (!c || !ms4) return EEMD_ERR_BAD_ARG;
  for (int k = 0; k < 4; ++k) ms4[k] = 0.0;
  if (c->last_phase_grid <= 0) return fail(c, EEMD_ERR_BAD_ARG, "last call ran without EEMD_FLAG_PHASES");
  std::vector<uint64_t> h((size_t)c->last_phase_grid * 4);
  HIPCHK(c, hipStreamSynchronize(c->stream), "phase sync");
  HIPCHK(c, hipMemcpy(h.data(), c->phase.p, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost),
         "phase copy");
  for (int g = 0; g < c->last_phase_grid; ++g)
    for (int k = 0; k < 4; ++k) ms4[k] += (double)h[(size_t)g * 4 + k];
  // wall_clock64 ticks at 100 MHz; report the mean over workgroups in ms
  for (int k = 0; k < 4; ++k) ms4[k] = ms4[k] / c->last_phase_grid / 1.0e5;
  return EEMD_OK;
}

}  // extern "C"

int eemd::ctx_device(const eemd_ctx* c) { return c ? c->device : 0; }
void* eemd::ctx_stream(eemd_ctx* c) { return c ? (void*)c->stream : nullptr; }
int eemd::ctx_fail(eemd_ctx* c, int code, const char* msg) { return fail(c, code, msg); }
void*& eemd::ctx_spectra(eemd_ctx* c) { return c->spectra; }
void eemd::ctx_spectra_dtor(eemd_ctx* c, void (*f)(void*)) { c->spectra_free = f; }
