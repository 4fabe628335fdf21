"""Learned boundary extension: data, training and evaluation.

No counterpart exists in the reference: its GPU kernels use natural ends, and its
CPU oracle reflects one extremum past each end (emd_ref.py:92-95).  Here an MLP
(64 -> 128 -> 128 -> 128 -> 16, GELU, BF16 parameters) predicts, at each end of
every sift, the next four extrema of both envelopes. It reads the 16 extrema
nearest to that end.  The sift kernel runs the model in-kernel
(``end_mode='learned'``) and falls back to the mirror rule when an envelope has
fewer than 16 extrema.

Training data come from the engine itself.  Long chirp/AM signals (16,384
samples, seed 3) are decomposed in FP64 with mirror ends.  Windows of 4,096
samples are cut from them, far from the long signals' own ends.  Each window
gives two examples per decomposition state, one per end:

* model input: the window's 16 extrema nearest to the end on both envelopes
  (for strict extrema, the window's interior extrema are exactly the long
  signal's extrema there);
* target: the long signal's next 4 extrema beyond that end.

The states used are the residual at the start of IMF levels 0..3 and the
finished IMFs 0..3.

Encoding, shared with the kernel (eemd4cuda_amd/csrc/eemd_mlp.h and
sift kernel ``learned_ends``).  Per end, distances are measured from the end
sample.

* Inputs: the gaps d_0 (nearest extremum to the end) and d_i (between
  neighbours), encoded as d / G - 1, where G is the mean of the 30
  inter-extremum gaps.  The values are encoded as (y - mu) / sd over the 32
  values.
* Outputs, per envelope: o[0] = log((d'_0 + 1) / G), where d'_0 is the first
  true extremum's distance beyond the end sample; o[1..3] = log(gap / G)
  between consecutive true extrema; o[4..7] = (y' - mu) / sd.  Outputs are
  clipped to [-4, 4].

Training runs on the GPU through the engine's own HIP kernels (eemd_mlp_*:
BF16 MFMA GEMMs, FP32 accumulation, Adam).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib, signals

CTX = 16
PRED = 4
N_IN = 64
N_OUT = 16
NPARAM = 43408


# ---------------------------------------------------------------------------
# data
# ---------------------------------------------------------------------------
def strict_extrema(h: np.ndarray):
    """Indices of strict local maxima and minima, time order (no endpoints)."""
    a, b, c = h[:-2], h[1:-1], h[2:]
    mx = np.nonzero((b > a) & (b > c))[0] + 1
    mn = np.nonzero((b < a) & (b < c))[0] + 1
    return mx, mn


def _end_examples(h, kx, kn, starts, W):
    """Input/target rows for every window start in `starts` of one state h
    (knots kx / kn of h), both ends.  Windows lacking 16 extrema inside or 4
    beyond an end on either envelope are skipped."""
    X, Y = [], []
    ks = (kx, kn)
    for a in starts:
        e = a + W - 1
        lo = [np.searchsorted(k, a + 1, "left") for k in ks]
        hi = [np.searchsorted(k, e - 1, "right") for k in ks]
        ok = all(hi[v] - lo[v] >= CTX and lo[v] >= PRED and len(ks[v]) - hi[v] >= PRED
                 for v in range(2))
        if not ok:
            continue
        for end in (0, 1):
            ls = np.zeros((2, CTX))
            ly = np.zeros((2, CTX))
            tt = np.zeros((2, PRED))
            ty = np.zeros((2, PRED))
            for v in range(2):
                k = ks[v]
                if end:
                    inner = k[hi[v] - 1 - np.arange(CTX)]
                    outer = k[hi[v] + np.arange(PRED)]
                    ls[v] = e - inner
                    tt[v] = outer - e
                else:
                    inner = k[lo[v] + np.arange(CTX)]
                    outer = k[lo[v] - 1 - np.arange(PRED)]
                    ls[v] = inner - a
                    tt[v] = a - outer
                ly[v] = h[inner]
                ty[v] = h[outer]
            gaps = np.diff(ls, axis=1)
            G = np.float32(gaps.astype(np.float32).sum() / np.float32(2 * (CTX - 1)))
            mu = np.float32(ly.astype(np.float32).sum() / np.float32(2 * CTX))
            sd = np.float32(max(np.sqrt(np.mean((ly.astype(np.float32) - mu) ** 2)), 1e-30))
            x = np.zeros(N_IN, np.float32)
            y = np.zeros(N_OUT, np.float32)
            for v in range(2):
                d = np.concatenate(([ls[v, 0]], gaps[v])).astype(np.float32)
                x[v * 32: v * 32 + 16] = d / G - 1.0
                x[v * 32 + 16: v * 32 + 32] = (ly[v].astype(np.float32) - mu) / sd
                tg = np.concatenate(([tt[v, 0] + 1.0], np.diff(tt[v])))
                y[v * 8: v * 8 + 4] = np.clip(np.log(tg / G), -4, 4)
                y[v * 8 + 4: v * 8 + 8] = np.clip((ty[v] - mu) / sd, -4, 4)
            X.append(x)
            Y.append(y)
    return X, Y


@dataclass
class Dataset:
    x: np.ndarray   # [n, 64] float32
    y: np.ndarray   # [n, 16] float32


def long_signal_states(xs: np.ndarray, max_imf: int = 4, engine=None):
    """FP64 EMD (mirror ends, 10 sifts) of each long signal on the engine:
    per signal the residuals at the start of levels 0..max_imf-1 and the
    IMFs 0..max_imf-1."""
    eng = engine or _lib.engine()
    n_sig, L = xs.shape
    prm = _lib.default_params()
    prm.fp64 = 1
    prm.end_mode = _lib.END_MODES["mirror"]
    prm.max_imf = max_imf
    prm.max_sifts = 10
    prm.block = 1024
    prm.halo = 30
    offs = np.arange(n_sig + 1, dtype=np.int64) * L
    out = eng.run_batch(xs.ravel(), offs, 1, np.zeros(n_sig), prm, member_imfs=True)
    imfs = out["member_imfs"].reshape(n_sig, max_imf, L)
    states = []
    for s in range(n_sig):
        r = xs[s].copy()
        st = []
        for m in range(int(out["nmode"][s, 0])):
            st.append(r.copy())
            st.append(imfs[s, m].copy())
            r = r - imfs[s, m]
        states.append(st)
    return states


def make_dataset(n_windows: int = 20000, W: int = 4096, L: int = 16384, windows_per_signal: int = 10,
                 seed: int = 3, engine=None) -> Dataset:
    """Training examples from n_windows windows (see the module docstring)."""
    rng = np.random.default_rng(seed)
    n_sig = (n_windows + windows_per_signal - 1) // windows_per_signal
    xs = np.stack([signals.chirp_am(L, rng) for _ in range(n_sig)])
    states = long_signal_states(xs, engine=engine)
    margin = 1024
    X, Y = [], []
    for s in range(n_sig):
        starts = rng.integers(margin, L - W - margin, windows_per_signal)
        for h in states[s]:
            kx, kn = strict_extrema(h)
            a, b = _end_examples(h, kx, kn, starts, W)
            X += a
            Y += b
    return Dataset(np.array(X, np.float32), np.array(Y, np.float32))


# ---------------------------------------------------------------------------
# training (eemd_mlp_* C ABI)
# ---------------------------------------------------------------------------
def _mlp_lib():
    L = _lib.load()
    if not hasattr(L, "_mlp_bound"):
        vp, i32, u64 = C.c_void_p, C.c_int32, C.c_uint64
        fp = C.POINTER(C.c_float)
        L.eemd_mlp_create.argtypes = [C.c_int, u64, C.POINTER(vp)]
        L.eemd_mlp_destroy.argtypes = [vp]
        L.eemd_mlp_last_error.argtypes = [vp]
        L.eemd_mlp_last_error.restype = C.c_char_p
        L.eemd_mlp_set.argtypes = [vp, vp]
        L.eemd_mlp_get.argtypes = [vp, vp]
        L.eemd_mlp_step.argtypes = [vp, vp, vp, i32, C.c_float, fp]
        L.eemd_mlp_grad.argtypes = [vp, vp, vp, i32, vp, fp]
        L.eemd_mlp_infer.argtypes = [vp, vp, i32, vp]
        L.eemd_set_boundary_model.argtypes = [vp, vp]
        for nm in ("eemd_mlp_create", "eemd_mlp_destroy", "eemd_mlp_set", "eemd_mlp_get",
                   "eemd_mlp_step", "eemd_mlp_grad", "eemd_mlp_infer", "eemd_set_boundary_model"):
            getattr(L, nm).restype = C.c_int
        L._mlp_bound = True
    return L


class MLP:
    """Training state of the boundary model on one GPU (eemd_mlp_*)."""

    def __init__(self, seed: int = 0, device: int = 0):
        self.L = _mlp_lib()
        h = C.c_void_p()
        rc = self.L.eemd_mlp_create(device, C.c_uint64(seed), C.byref(h))
        if rc != _lib.EEMD_OK:
            raise _lib.EEMDBackendError(f"eemd_mlp_create failed with code {rc}")
        self.h = h

    def _check(self, rc, what):
        if rc == _lib.EEMD_OK:
            return
        msg = (self.L.eemd_mlp_last_error(self.h) or b"").decode(errors="replace")
        if rc == _lib.EEMD_ERR_BAD_ARG:
            raise ValueError(f"{what}: {msg}")
        raise _lib.EEMDBackendError(f"{what} failed (code {rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.L.eemd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(NPARAM, np.float32)
        self._check(self.L.eemd_mlp_get(self.h, out.ctypes.data_as(C.c_void_p)), "eemd_mlp_get")
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != NPARAM:
            raise ValueError("wrong parameter count")
        self._check(self.L.eemd_mlp_set(self.h, p.ctypes.data_as(C.c_void_p)), "eemd_mlp_set")

    def step(self, x_ptr: int, y_ptr: int, batch: int, lr: float) -> float:
        loss = C.c_float()
        self._check(self.L.eemd_mlp_step(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), batch,
                                         C.c_float(lr), C.byref(loss)), "eemd_mlp_step")
        return loss.value

    def grad(self, x_ptr: int, y_ptr: int, batch: int):
        g = np.zeros(NPARAM, np.float32)
        loss = C.c_float()
        self._check(self.L.eemd_mlp_grad(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), batch,
                                         g.ctypes.data_as(C.c_void_p), C.byref(loss)),
                    "eemd_mlp_grad")
        return g, loss.value

    def infer(self, x_ptr: int, batch: int, y_ptr: int):
        self._check(self.L.eemd_mlp_infer(self.h, C.c_void_p(x_ptr), batch, C.c_void_p(y_ptr)),
                    "eemd_mlp_infer")


def train(ds: Dataset, epochs: int = 20, lr: float = 1e-3, batch: int = 256, seed: int = 0,
          val_frac: float = 0.1, log=None):
    """Adam on the mean squared error; returns (params, history)."""
    import torch
    rng = np.random.default_rng(seed)
    n = len(ds.x)
    perm = rng.permutation(n)
    nval = max(batch, (int(n * val_frac) // batch) * batch)
    val, tr = perm[:nval], perm[nval:]
    dev = torch.device("cuda", 0)
    xt = torch.from_numpy(ds.x).to(dev)
    yt = torch.from_numpy(ds.y).to(dev)
    xv, yv = xt[val].contiguous(), yt[val].contiguous()
    net = MLP(seed=seed)
    hist = []
    nb = len(tr) // batch
    for ep in range(epochs):
        order = torch.from_numpy(rng.permutation(tr)).to(dev)
        tl = 0.0
        for b in range(nb):
            idx = order[b * batch:(b + 1) * batch]
            xb = xt.index_select(0, idx).contiguous()
            yb = yt.index_select(0, idx).contiguous()
            torch.cuda.synchronize()
            tl += net.step(xb.data_ptr(), yb.data_ptr(), batch, lr)
        vl = 0.0
        for b in range(nval // batch):
            _, l_ = net.grad(xv[b * batch:].data_ptr(), yv[b * batch:].data_ptr(), batch)
            vl += l_
        hist.append((ep, tl / max(nb, 1), vl / max(nval // batch, 1)))
        if log:
            log("epoch %d train %.4f val %.4f" % hist[-1])
    return net.params(), hist


def default_model_path() -> str:
    """The model shipped with the package (trained by ``main`` with its
    defaults: 20,000 windows, 20 epochs, lr 1e-3, batch 256, seed 3)."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "boundary_model.npz")


def save(path: str, params: np.ndarray, **meta):
    np.savez(path, params=np.asarray(params, np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def load(path: str) -> np.ndarray:
    d = np.load(path, allow_pickle=False)
    p = np.asarray(d["params"], np.float32)
    if p.size != NPARAM:
        raise ValueError("not a boundary model: wrong parameter count")
    return p


# ---------------------------------------------------------------------------
# engine use
# ---------------------------------------------------------------------------
def install(params, engine=None):
    """Make `params` the boundary model of the engine (end_mode 'learned')."""
    eng = engine or _lib.engine()
    L = _mlp_lib()
    p = np.ascontiguousarray(params, np.float32)
    if p.size != NPARAM:
        raise ValueError("wrong parameter count")
    eng._check(L.eemd_set_boundary_model(eng.h, p.ctypes.data_as(C.c_void_p)),
               "eemd_set_boundary_model")


def run(x, offsets, E, sigmas, prm, model=None, noise=None, engine=None, **kw):
    """eemd_run_batch with end_mode 'learned' and `model` (parameters or an
    .npz path; None: the packaged model) installed."""
    eng = engine or _lib.engine()
    if model is None:
        model = default_model_path()
    if isinstance(model, (str, os.PathLike)):
        model = load(model)
    install(model, eng)
    prm.end_mode = _lib.END_MODES["learned"]
    return eng.run_batch(x, offsets, E, sigmas, prm, noise=noise, **kw)


def edge_error(params, n_windows: int = 256, W: int = 4096, L: int = 16384, seed: int = 5,
               max_imf: int = 3, fp64: int = 1, E: int = 1, eps: float = 0.0, engine=None):
    """Relative L2 error of window IMFs against the long signal's IMFs over
    the edge regions (first and last ceil(W/16) samples), mirror vs learned.

    Windows come from fresh long signals (seed `seed`), four per signal.  The
    truth is the long signal's FP64 decomposition (mirror ends, 10 sifts)
    restricted to the window.  Each window is decomposed alone (10 sifts;
    FP64 or FP32 by `fp64`) with mirror and with learned ends.  With E > 1
    both are EEMDs with E members: member e of a window receives the long
    signal's member-e noise restricted to the window (noise std eps * std of
    the long signal), so the window ensemble differs from the truth only by
    what its ends do.  Returns per-IMF errors (aggregate over the windows and
    the median window) for both end modes and their ratios.
    """
    eng = engine or _lib.engine()
    rng = np.random.default_rng(seed)
    n_sig = max(1, n_windows // 4)
    xs = np.stack([signals.chirp_am(L, rng) for _ in range(n_sig)])
    noisy = E > 1 or eps > 0.0
    z = rng.standard_normal((n_sig, E, L)) if noisy else None
    sig_long = eps * np.std(xs, axis=1) if noisy else np.zeros(n_sig)
    prm = _lib.default_params()
    prm.fp64 = 1
    prm.end_mode = _lib.END_MODES["mirror"]
    prm.max_imf = max_imf
    prm.max_sifts = 10
    offs = np.arange(n_sig + 1, dtype=np.int64) * L
    full = eng.run_batch(xs.ravel(), offs, E, sig_long, prm,
                         noise=None if z is None else z.ravel())
    imf_long = full["avg"].reshape(n_sig, max_imf, L)
    starts = rng.integers(1024, L - W - 1024, (n_sig, 4))
    wins = np.stack([xs[s, a:a + W] for s in range(n_sig) for a in starts[s]])
    truth = np.stack([imf_long[s, :, a:a + W] for s in range(n_sig) for a in starts[s]])
    wz = None
    wsig = np.zeros(len(wins))
    if noisy:
        wz = np.stack([z[s, :, a:a + W] for s in range(n_sig) for a in starts[s]])
        wsig = np.repeat(sig_long, 4)
    ew = -(-W // 16)
    edge = np.zeros(W, bool)
    edge[:ew] = True
    edge[-ew:] = True
    res = {}
    woffs = np.arange(len(wins) + 1, dtype=np.int64) * W
    for mode in ("mirror", "learned"):
        p = _lib.default_params()
        p.fp64 = fp64
        p.max_imf = max_imf
        p.max_sifts = 10
        if not fp64:
            p.halo = 14
        if mode == "learned":
            install(params, eng)
        p.end_mode = _lib.END_MODES[mode]
        o = eng.run_batch(wins.ravel(), woffs, E, wsig, p,
                          noise=None if wz is None else wz.ravel())
        got = o["avg"].reshape(len(wins), max_imf, W)
        errs, med = [], []
        for m in range(max_imf):
            d2 = np.sum((got[:, m, edge] - truth[:, m, edge]) ** 2, axis=1)
            t2 = np.sum(truth[:, m, edge] ** 2, axis=1)
            errs.append(float(np.sqrt(d2.sum() / max(t2.sum(), 1e-300))))
            med.append(float(np.median(np.sqrt(d2 / np.maximum(t2, 1e-300)))))
        res[mode] = errs
        res[mode + "_median_window"] = med
    res["ratio"] = [l / m if m > 0 else float("nan") for l, m in zip(res["learned"], res["mirror"])]
    res["ratio_median_window"] = [l / m if m > 0 else float("nan") for l, m in
                                  zip(res["learned_median_window"], res["mirror_median_window"])]
    res["edge_samples"] = ew
    res["windows"] = len(wins)
    res["members"] = E
    res["fp"] = "f64" if fp64 else "f32"
    return res


def main(argv=None):
    """python -m eemd4cuda_amd.learned --out model.npz [--windows 20000] [--epochs 20]"""
    import argparse
    import json
    import time
    ap = argparse.ArgumentParser(description="train the boundary-extension model on the GPU")
    ap.add_argument("--out", required=True)
    ap.add_argument("--windows", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--eval-windows", type=int, default=256)
    a = ap.parse_args(argv)
    t0 = time.time()
    ds = make_dataset(a.windows, seed=a.seed)
    print("dataset: %d examples from %d windows (%.1f s)" % (len(ds.x), a.windows, time.time() - t0),
          flush=True)
    t1 = time.time()
    params, hist = train(ds, epochs=a.epochs, lr=a.lr, batch=a.batch,
                         log=lambda m: print(m, flush=True))
    print("training %.1f s" % (time.time() - t1), flush=True)
    ev = edge_error(params, n_windows=a.eval_windows)
    print("edge error:", json.dumps(ev), flush=True)
    save(a.out, params, windows=a.windows, epochs=a.epochs, lr=a.lr, batch=a.batch,
         seed=a.seed, history=np.array(hist), edge_mirror=np.array(ev["mirror"]),
         edge_learned=np.array(ev["learned"]),
         edge_mirror_median=np.array(ev["mirror_median_window"]),
         edge_learned_median=np.array(ev["learned_median_window"]))


if __name__ == "__main__":
    main()
