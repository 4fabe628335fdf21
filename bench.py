"""Benchmark of the MI355X EEMD engine.

Default workload (one MI355X): plain EEMD of one mono 48 kHz record of 2**20
samples (six prime tones, an AM tone, a linear chirp, pink noise at -30 dB,
seed 1), 256 ensemble members with noise std 0.2*std(x), 10 IMFs x 10 sifts,
mirror ends, decay-seamed block-Thomas with block 1024.  One "step" is one
complete EEMD of that record (all members, all IMFs, ensemble mean and
residual) with the input already resident in HBM and the noise generated
in-kernel (Philox).  The headline runs in FP64, the reference's precision,
bit-exact with the oracle; FP32 is measured alongside with its per-IMF error
against FP64.

Metric: ensemble sample-sifts/s = sum over members of N x sifts executed,
divided by the wall time of the timed steps.  roofline.achieved prices the
persistent sift kernel at 10.4 algorithmic bytes per sample-sift in FP32
(19.2 in FP64) against the 8 TB/s HBM3E peak.  cpu_baseline times the
reference's CPU path, restated in numpy (oracle/emd_np.py), on a bounded
sample of members of the same workload, one member per host core.

    python bench.py [--steps 5] [--warmup 1] [--workload audio_mix|three_tone|uneven|chirp_scale]
                    [--only f64|f32 [--end mirror|learned]] [--no-legs] [--no-cpu]

Besides the headline, the default run reports two bounded legs of the
metric's own comparison: plain per-signal launches against the batched,
graph-replayed and autotuned launch on 128 signals of the uneven batch
(FP64 and FP32), and mirror against learned ends on 512 signals of the
4096 x 65536 chirp/AM set (FP32: throughput, cost, FP32 error per IMF against
FP64 on 32 of them, and the edge-region IMF error of 256 held-out windows).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def _log(msg):
    """Progress on stderr (stdout carries only the JSON result line)."""
    print("[bench %.0fs] %s" % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)


_T0 = time.perf_counter()
BYTES_PER_SAMPLE_SIFT = {"f32": 10.4, "f64": 19.2}


def workload(name: str):
    from eemd4cuda_amd import signals
    if name == "audio_mix":
        x = signals.audio_mix(1 << 20, seed=1)
        return dict(xs=[x], E=256, max_imf=10, max_sifts=10, B=1024,
                    label="audio_mix N=2^20 E=256 10 IMFs x 10 sifts mirror B=1024")
    if name == "three_tone":
        return dict(xs=[signals.three_tone(4096)], E=64, max_imf=8, max_sifts=10, B=1024,
                    label="three_tone N=4096 E=64 8 IMFs x 10 sifts mirror")
    if name == "uneven":
        xs = signals.uneven_batch(512, seed=2)
        return dict(xs=xs, E=100, max_imf=8, max_sifts=10, B=1024,
                    label="uneven 512 signals N in [2^12,2^18] E=100 8 IMFs x 10 sifts mirror")
    if name == "chirp_scale":
        xs = list(signals.chirp_am_batch(4096, 65536, seed=4))
        return dict(xs=xs, E=128, max_imf=10, max_sifts=10, B=1024,
                    label="chirp_am 4096 x 65536 E=128 10 IMFs x 10 sifts mirror")
    raise SystemExit("unknown workload " + name)


class DeviceRun:
    """Device-resident inputs/outputs for repeated eemd_run_batch_dev calls."""

    def __init__(self, wl, fp: str, halo: int, threads: int = 0, mpt: int = 0, wave_windows: int = 0):
        import torch
        from eemd4cuda_amd import _lib
        self.torch = torch
        self.eng = _lib.engine()
        xs = wl["xs"]
        self.lens = np.array([len(v) for v in xs], dtype=np.int64)
        self.offs = np.concatenate(([0], np.cumsum(self.lens))).astype(np.int64)
        self.E = wl["E"]
        self.sig = np.array([0.2 * np.std(v) for v in xs])
        self.mi = wl["max_imf"]
        dev = torch.device("cuda", 0)
        ntot = int(self.offs[-1])
        nsig = len(xs)
        self.x = torch.from_numpy(np.concatenate(xs)).to(dev)
        self.avg = torch.zeros(self.mi * ntot, dtype=torch.float64, device=dev)
        self.resid = torch.zeros(ntot, dtype=torch.float64, device=dev)
        self.n_modes = torch.zeros(nsig, dtype=torch.int32, device=dev)
        self.nmode = torch.zeros(nsig * self.E, dtype=torch.int32, device=dev)
        self.nsift = torch.zeros(nsig * self.E, dtype=torch.int32, device=dev)
        p = _lib.default_params()
        p.fp64 = 1 if fp == "f64" else 0
        p.end_mode = _lib.END_MODES["mirror"]
        p.max_imf = self.mi
        p.max_sifts = wl["max_sifts"]
        p.block = wl["B"]
        p.halo = halo
        p.threads = threads
        p.members_per_task = mpt
        p.wave_windows = wave_windows
        p.seed = 1
        self.p = p
        torch.cuda.synchronize()

    def step(self, phases=False, graph=False, tune_block=False):
        from eemd4cuda_amd import _lib
        self.p.flags = ((_lib.FLAG_PHASES if phases else 0) | (_lib.FLAG_GRAPH if graph else 0)
                        | (_lib.FLAG_TUNE_BLOCK if tune_block else 0))
        self.eng.run_batch_dev(self.x.data_ptr(), self.offs, self.E, self.sig, self.p,
                               self.avg.data_ptr(), self.resid.data_ptr(),
                               self.n_modes.data_ptr(), self.nmode.data_ptr(),
                               self.nsift.data_ptr())
        return self.eng.last_timing()

    def step_per_signal(self):
        """The plain comparator: one engine call per signal, same arithmetic."""
        from eemd4cuda_amd import _lib
        self.p.flags = 0
        for s in range(len(self.lens)):
            self.p.signal0 = s
            o, n = int(self.offs[s]), int(self.lens[s])
            self.eng.run_batch_dev(self.x.data_ptr() + 8 * o, np.array([0, n]), self.E,
                                   self.sig[s:s + 1], self.p,
                                   self.avg.data_ptr() + 8 * self.mi * o,
                                   self.resid.data_ptr() + 8 * o,
                                   self.n_modes.data_ptr() + 4 * s,
                                   self.nmode.data_ptr() + 4 * s * self.E,
                                   self.nsift.data_ptr() + 4 * s * self.E)
        self.p.signal0 = 0

    def autotune(self, reps=1, tune_block=False):
        from eemd4cuda_amd import _lib
        self.p.flags = _lib.FLAG_TUNE_BLOCK if tune_block else 0
        best, ms = self.eng.autotune(self.x.data_ptr(), self.offs, self.E, self.sig, self.p,
                                     reps=reps)
        self.p.flags = 0
        return best, ms

    def sample_sifts(self) -> int:
        ns = self.nsift.cpu().numpy().reshape(len(self.lens), self.E).astype(np.int64)
        return int((ns * self.lens[:, None]).sum())

    def imfs(self):
        a = self.avg.cpu().numpy()
        nm = self.n_modes.cpu().numpy()
        out = []
        for s in range(len(self.lens)):
            blk = a[self.mi * self.offs[s]: self.mi * self.offs[s + 1]].reshape(self.mi, self.lens[s])
            out.append(blk[:nm[s]])
        return out


def time_mode(wl, fp, steps, warmup, halo, threads=0, mpt=0, wave_windows=0, end="mirror"):
    import torch
    from eemd4cuda_amd import _lib
    r = DeviceRun(wl, fp, halo, threads=threads, mpt=mpt, wave_windows=wave_windows)
    r.p.end_mode = _lib.END_MODES[end]
    for _ in range(warmup):
        r.step()
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        tm = r.step()
        kms.append(tm["sift_ms"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ss = r.sample_sifts()
    ms = wall / steps * 1e3
    k_s = float(np.mean(kms)) / 1e3
    achieved = ss * BYTES_PER_SAMPLE_SIFT[fp] / k_s / 1e9
    return dict(value=ss / (wall / steps), ms_per_step=ms, sample_sifts=ss,
                sift_kernel_ms=float(np.mean(kms)), mean_kernel_ms=tm["mean_ms"],
                threads=tm["threads"], grid=tm["grid"], ntasks=tm["ntasks"],
                roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=achieved / HBM_PEAK_GBS, traffic=None)), r


def learned_leg(wl, wl_err, steps, warmup, halo):
    """Learned against mirror ends with the packaged model, FP32: throughput
    of both on `wl` and the cost of the learned ends; the FP32 error per
    ensemble-mean IMF against FP64 (both with learned ends, same noise) on
    `wl_err`; and the edge-region IMF error of mirror vs learned ends on 256
    held-out chirp/AM windows (FP64, E=1)."""
    import torch
    from eemd4cuda_amd import _lib, learned
    params = learned.load(learned.default_model_path())
    learned.install(params)
    out = {"workload": wl["label"], "model": "eemd4cuda_amd/boundary_model.npz"}
    vals = {}
    for mode in ("mirror", "learned"):
        r = DeviceRun(wl, "f32", halo)
        r.p.end_mode = _lib.END_MODES[mode]
        for _ in range(warmup):
            r.step()
        torch.cuda.synchronize()
        t = _timed(r, steps, r.step)
        tm = r.eng.last_timing()
        ss = r.sample_sifts()
        vals[mode] = ss / t
        out[mode] = dict(value=ss / t, ms_per_step=t * 1e3, sift_kernel_ms=tm["sift_ms"],
                         roofline_frac=ss * BYTES_PER_SAMPLE_SIFT["f32"] / (tm["sift_ms"] / 1e3)
                         / 1e9 / HBM_PEAK_GBS)
        del r
    out["cost_vs_mirror"] = 1.0 - vals["learned"] / vals["mirror"]
    runs = {}
    for fp, h in (("f64", 30), ("f32", halo)):
        r = DeviceRun(wl_err, fp, h)
        r.p.end_mode = _lib.END_MODES["learned"]
        r.step()
        runs[fp] = r
    e_imf, e_sig = imf_errors(runs["f32"], runs["f64"], wl_err["xs"])
    out["fp32_error_workload"] = wl_err["label"] + " learned ends"
    out["fp32_imf_rel_l2_vs_fp64"] = e_imf
    out["fp32_imf_l2_vs_fp64_rel_to_signal"] = e_sig
    del runs
    out["edge_error"] = learned.edge_error(params, n_windows=256)
    return out


def _timed(r, steps, fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def comparators(wl, fp, steps, halo, tune_block):
    """Plain per-signal launches vs the batched persistent launch, then with
    hipGraph replay and with the autotuned configuration (workgroup size
    including one wave per member, members per task, FP32 wave-solve
    threshold and, with tune_block, the block-Thomas block size).  Every
    launch decomposes the same members (a per-signal call keys its Philox
    noise by the signal's index in the batch).  Reported: sift and IMF counts
    equal to the batched run, and the largest relative L2 difference of an
    ensemble-mean IMF (FP64 without a block change: re-association only)."""
    r = DeviceRun(wl, fp, halo)
    r.step()
    ss = r.sample_sifts()

    def snap():
        return (r.avg.cpu().numpy().copy(), r.nsift.cpu().numpy().copy(),
                r.nmode.cpu().numpy().copy())

    def compare(ref):
        a, ns, nm = snap()
        worst = 0.0
        for s_ in range(len(r.lens)):
            o, n = int(r.offs[s_]), int(r.lens[s_])
            A = a[r.mi * o: r.mi * (o + n)].reshape(r.mi, n)
            B = ref[0][r.mi * o: r.mi * (o + n)].reshape(r.mi, n)
            for m in range(r.mi):
                nb = float(np.linalg.norm(B[m]))
                if nb > 0:
                    worst = max(worst, float(np.linalg.norm(A[m] - B[m])) / nb)
        return dict(counts_equal=bool(np.array_equal(ns, ref[1]) and np.array_equal(nm, ref[2])),
                    max_imf_rel_l2=worst)
    ref = snap()
    out = {}
    t_plain = _timed(r, steps, r.step_per_signal)
    out["plain_per_signal"] = dict(value=ss / t_plain, ms_per_step=t_plain * 1e3,
                                   launches=len(r.lens), vs_batched=compare(ref))
    t_b = _timed(r, steps, r.step)
    tm = r.eng.last_timing()
    out["batched"] = dict(value=ss / t_b, ms_per_step=t_b * 1e3, threads=tm["threads"])
    r.step(graph=True)
    t_g = _timed(r, steps, lambda: r.step(graph=True))
    out["batched_graph"] = dict(value=ss / t_g, ms_per_step=t_g * 1e3)
    t0 = time.perf_counter()
    best, _ = r.autotune(reps=1, tune_block=tune_block)
    tune_s = time.perf_counter() - t0
    # automatic parameters now resolve to the tuned choice
    r.step(graph=True, tune_block=tune_block)
    t_t = _timed(r, steps, lambda: r.step(graph=True, tune_block=tune_block))
    tm = r.eng.last_timing()
    out["batched_graph_autotuned"] = dict(
        value=ss / t_t, ms_per_step=t_t * 1e3, tune_s=tune_s, vs_batched=compare(ref),
        choice=dict(threads=best.threads, members_per_task=best.members_per_task,
                    wave_windows=best.wave_windows, block=best.block))
    out["batched_vs_plain"] = out["batched_graph_autotuned"]["value"] / out["plain_per_signal"]["value"]
    out["sample_sifts_per_step"] = ss
    # graph replays are timed as a whole on the engine's stream
    out["roofline_frac_batched"] = (ss * BYTES_PER_SAMPLE_SIFT[fp] / (tm["sift_ms"] / 1e3) / 1e9
                                    / HBM_PEAK_GBS)
    return out


def reference_sweep_leg():
    """The reference's own GPU benchmark (published wall times:
    Benchmark_Thomas_vs_Decay-Seamed_Block-Thomas.md:70-73, 150-153): EEMD of
    the 441,000-sample 10-octave sweep (refsignal.py,
    emd_par.py:307-318) through decompose(method='gpu') with the reference's
    settings: numpy noise default_rng(0), eps 0.2, SD stop tau 0.25 (at most
    50 sifts), 16 IMFs, natural ends, block-Thomas B=250 / S=30, FP64.  Wall
    time of the whole call, host noise and PCIe copies included, as the
    reference measures it."""
    from eemd4cuda_amd import emd_par, signals
    x = signals.reference_sweep()
    out = {"signal": "reference 10-octave sweep, N=441000", "published_wall_s": {
        "rtx4070_laptop": {"8": 0.83, "32": 1.77, "64": 2.54, "128": 11.08},
        "gb10": {"8": 1.73, "32": 3.87, "64": 4.57, "128": 6.22}}}
    kw = dict(method="gpu", eps=0.2, tau=0.25, max_sifts=50, max_imf=16, end="natural", B=250,
              S=30, seed=0)
    emd_par.decompose(x, E=8, **kw)   # warm-up (first use of the context's buffers)
    for E in (8, 32, 64, 128):
        t0 = time.perf_counter()
        r = emd_par.decompose(x, E=E, **kw)
        wall = time.perf_counter() - t0
        out["E%d" % E] = dict(wall_s=wall, n_modes=int(r.n_modes),
                              total_sifts=int(r.stats["total_sifts"]),
                              sample_sifts_per_s=r.stats["sample_sifts"] / wall)
        # the same with the noise drawn in-kernel (Philox) instead of on the host
        t0 = time.perf_counter()
        r = emd_par.decompose(x, E=E, noise="philox", **kw)
        wall = time.perf_counter() - t0
        out["E%d_philox" % E] = dict(wall_s=wall, total_sifts=int(r.stats["total_sifts"]),
                                     sample_sifts_per_s=r.stats["sample_sifts"] / wall)
    return out


def cpu_baseline(wl, budget_cores: int):
    """The reference's CPU path (numpy restatement) on one member per core."""
    from concurrent.futures import ProcessPoolExecutor

    from oracle import emd_np, orc
    cores = max(1, min(budget_cores, os.cpu_count() or 1))
    x = wl["xs"][0]
    sigma = 0.2 * np.std(x)
    tasks = [(x, orc.normals(1, 0, e, len(x)), sigma, 0.0, wl["max_sifts"], wl["max_imf"], 4,
              "mirror") for e in range(cores)]
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=cores) as ex:
        res = list(ex.map(emd_np.eemd_member, tasks))
    wall = time.perf_counter() - t0
    ss = sum(len(x) * r[1]["total_sifts"] for r in res)
    return dict(value=ss / wall, unit="sample-sifts/s", cores=cores, kind="port",
                sample="%d members (one per core) of the same record, full %d IMFs x %d sifts, "
                       "numpy restatement of emd_ref (oracle/emd_np.py), ProcessPoolExecutor; "
                       "wall %.1f s" % (cores, wl["max_imf"], wl["max_sifts"], wall))


def cpu_baseline_c(wl, budget_cores: int):
    """The C oracle (OpenMP over members): a stronger CPU bar."""
    from oracle import orc
    cores = max(1, min(budget_cores, os.cpu_count() or 1))
    x = wl["xs"][0]
    E = cores
    z = np.stack([orc.normals(1, 0, e, len(x)) for e in range(E)])
    t0 = time.perf_counter()
    o = orc.eemd(x, z, 0.2 * np.std(x), tau=0.0, max_sifts=wl["max_sifts"], max_imf=wl["max_imf"],
                 end="mirror", B=0, nthreads=cores)
    wall = time.perf_counter() - t0
    ss = int(o["nsift"].astype(np.int64).sum()) * len(x)
    return dict(value=ss / wall, unit="sample-sifts/s", cores=cores, kind="port",
                sample="%d members, C oracle -O2 OpenMP, exact Thomas; wall %.1f s" % (E, wall))


def copy_probe(mib: int):
    """A plain copy of known size: its FETCH_SIZE / WRITE_SIZE readings calibrate the counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "launches": 3}), flush=True)


def uneven_subset(n: int):
    """The first n signals of the 512-signal uneven batch (seed 2)."""
    from eemd4cuda_amd import signals
    xs = signals.uneven_batch(512, seed=2)[:n]
    return dict(xs=xs, E=100, max_imf=8, max_sifts=10, B=1024,
                label="uneven first %d of 512 signals N in [2^12,2^18] E=100 8 IMFs x 10 sifts "
                      "mirror" % n)


def chirp_subset(n: int, E: int = 128):
    """The first n signals of the 4096 x 65536 chirp/AM set (seed 4)."""
    from eemd4cuda_amd import signals
    xs = list(signals.chirp_am_batch(n, 65536, seed=4))
    return dict(xs=xs, E=E, max_imf=10, max_sifts=10, B=1024,
                label="chirp_am first %d of 4096 x 65536 E=%d 10 IMFs x 10 sifts" % (n, E))


def imf_errors(r_a, r_b, xs):
    """Worst signal's relative L2 error per ensemble-mean IMF of run a
    against run b (same noise), relative to the IMF and to the signal."""
    ia, ib = r_a.imfs(), r_b.imfs()
    errs, errs_x = [], []
    for s_, (a_, b_) in enumerate(zip(ia, ib)):
        m = min(len(a_), len(b_))
        xn = max(float(np.linalg.norm(xs[s_])), 1e-300)
        errs.append([float(np.linalg.norm(a_[k] - b_[k]) / max(np.linalg.norm(b_[k]), 1e-300))
                     for k in range(m)])
        errs_x.append([float(np.linalg.norm(a_[k] - b_[k]) / xn) for k in range(m)])

    def worst(e):
        return e[0] if len(e) == 1 else \
            [float(np.max([v[k] for v in e if len(v) > k])) for k in range(max(map(len, e)))]
    return worst(errs), worst(errs_x)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="audio_mix")
    ap.add_argument("--only", choices=("f64", "f32"), default=None,
                    help="time only this precision on the workload (no legs, no CPU baseline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="skip the comparator and learned legs")
    ap.add_argument("--cpu-cores", type=int, default=16)
    ap.add_argument("--phases", action="store_true", help="report per-phase kernel time")
    ap.add_argument("--block", type=int, default=0, help="override the block-Thomas block size")
    ap.add_argument("--halo32", type=int, default=14)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--mpt", type=int, default=0, help="members per task (0: automatic)")
    ap.add_argument("--E", type=int, default=0, help="override the ensemble size")
    ap.add_argument("--max-imf", type=int, default=0, help="override the IMF count")
    ap.add_argument("--wave-windows", type=int, default=0, help="wave-parallel solve threshold")
    ap.add_argument("--end", choices=("mirror", "learned"), default="mirror",
                    help="end extension of an --only run (learned: the packaged model)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (scales the "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the EEMD engine runs on a single GPU here "
              "(--gpus 1).", file=sys.stderr)
        sys.exit(2)

    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    _log("building workload %s" % a.workload)
    wl = workload(a.workload)
    if a.block:
        wl["B"] = a.block
    if a.E:
        wl["E"] = a.E
    if a.max_imf:
        wl["max_imf"] = a.max_imf
    halo = {"f64": 30, "f32": a.halo32}
    if a.only:
        if a.end == "learned":
            from eemd4cuda_amd import learned
            learned.install(learned.load(learned.default_model_path()))
        res, run = time_mode(wl, a.only, a.steps, a.warmup, halo=halo[a.only], threads=a.threads,
                             mpt=a.mpt, wave_windows=a.wave_windows, end=a.end)
        label = wl["label"] if a.end == "mirror" else wl["label"].replace("mirror", "learned ends")
        line = dict(fp=a.only, workload=label, end=a.end,
                    **{k: res[k] for k in ("value", "ms_per_step", "sift_kernel_ms",
                                           "threads", "grid", "roofline")})
        if a.phases:
            run.step(phases=True)
            line["phases_ms"] = dict(zip(("solve", "pass", "fixup", "other"),
                                         run.eng.last_phases()))
        print(json.dumps(line), flush=True)
        return
    _log("timing f64 (headline)")
    f64, r64 = time_mode(wl, "f64", a.steps, a.warmup, halo=30, threads=a.threads, mpt=a.mpt)
    _log("f64 %.1f ms per step" % f64["ms_per_step"])
    out = {
        "metric": "ensemble sample-sifts/s",
        "value": f64["value"],
        "unit": "sample-sifts/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": f64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl["label"], "fp": "f64 (bit-exact with the oracle)",
                   "halo": 30, "block": wl["B"], "threads": f64["threads"], "grid": f64["grid"],
                   "tasks": f64["ntasks"], "noise": "philox in-kernel",
                   "sample_sifts_per_step": f64["sample_sifts"]},
        "roofline": f64["roofline"],
        "sift_kernel_ms": f64["sift_kernel_ms"],
        "mean_kernel_ms": f64["mean_kernel_ms"],
    }
    if a.phases:
        r64.step(phases=True)
        out["phases_ms_f64"] = dict(zip(("solve", "pass", "fixup", "other"), r64.eng.last_phases()))
    _log("timing f32")
    f32, r32 = time_mode(wl, "f32", a.steps, a.warmup, halo=a.halo32, threads=a.threads,
                         mpt=a.mpt, wave_windows=a.wave_windows)
    e_imf, e_sig = imf_errors(r32, r64, wl["xs"])
    out["fp32"] = {"value": f32["value"], "ms_per_step": f32["ms_per_step"],
                   "sift_kernel_ms": f32["sift_kernel_ms"], "roofline": f32["roofline"],
                   "halo": a.halo32, "threads": f32["threads"],
                   # FP32 error per ensemble-mean IMF against the FP64 run (same
                   # noise), relative to the IMF and to the signal
                   "imf_rel_l2_vs_fp64": e_imf, "imf_l2_vs_fp64_rel_to_signal": e_sig}
    if a.phases:
        r32.step(phases=True)
        out["fp32"]["phases_ms"] = dict(zip(("solve", "pass", "fixup", "other"),
                                            r32.eng.last_phases()))
    del r32, r64
    prof = ROOT / "profiles" / ("traffic_%s.json" % a.workload)
    if prof.exists():
        tr = json.loads(prof.read_text())
        out["roofline"]["traffic"] = tr.get("f64_bytes_per_launch")
        out["fp32"]["roofline"]["traffic"] = tr.get("f32_bytes_per_launch")
    if not a.no_legs and a.workload == "uneven":
        _log("leg: batched vs plain on the whole uneven batch")
        # the block size is tuned in FP64 only, where it moves the ensemble
        # means by rounding (FP32 decompositions change with the block size
        # well beyond that, so its tuned run keeps the workload's block)
        out["batched_vs_plain"] = {"workload": wl["label"],
                                   "f64": comparators(wl, "f64", 1, 30, tune_block=True),
                                   "f32": comparators(wl, "f32", 1, a.halo32, tune_block=False)}
    if not a.no_legs and a.workload == "chirp_scale":
        _log("leg: learned vs mirror ends on the whole chirp/AM set")
        out["learned_vs_mirror"] = learned_leg(wl, chirp_subset(32), a.steps, a.warmup, a.halo32)
    if not a.no_legs and a.workload == "audio_mix":
        _log("leg: batched vs plain on 128 uneven signals")
        sub = uneven_subset(128)
        out["batched_vs_plain"] = {"workload": sub["label"],
                                   "f64": comparators(sub, "f64", 1, 30, tune_block=True),
                                   "f32": comparators(sub, "f32", 1, a.halo32, tune_block=False)}
        _log("leg: learned vs mirror ends on 512 chirp/AM signals")
        out["learned_vs_mirror"] = learned_leg(chirp_subset(512), chirp_subset(32), 1, 1,
                                               a.halo32)
    if not a.no_legs and a.workload == "audio_mix":
        _log("leg: the reference's own benchmark (441,000-sample sweep)")
        out["reference_benchmark"] = reference_sweep_leg()
    if not a.no_cpu and len(wl["xs"]) >= 1:
        _log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_cores)
        if a.workload in ("audio_mix", "three_tone"):
            out["cpu_baseline_c"] = cpu_baseline_c(wl, a.cpu_cores)
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
