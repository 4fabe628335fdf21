"""GPU entry points with the reference's emd_gpu signatures, on the MI355X engine.

* :func:`gpu_eemd_full`  <- emd_gpu_v5_dsbt.py:354-384 (emd_gpu.py:301-334)
* :func:`gpu_eemd`       <- emd_gpu_v5_dsbt.py:387-410 (emd_gpu.py:338-358)
* :func:`gpu_emd_single` <- emd_gpu_v5_dsbt.py:413-433 (emd_gpu.py:361-377)

Semantics follow the reference's CPU oracle emd_ref.emd_1d (the reference's
numba kernels differ from it in three places, see emd_gpu.py:167, 191, 222):
the sift loop stops when h has no maximum or no minimum, a one-knot envelope
is the constant knot value, and the IMF loop stops on nmax+nmin < min_extrema.
`end` defaults to 'natural' like the reference kernels; 'mirror' applies
emd_ref.py:92-95.  The numba-specific knobs K and nthreads are accepted and
ignored (the MI355X engine sizes its own scratch and workgroups); B and S are
the decay-seamed block-Thomas block and halo, EXACT=True selects exact
Thomas.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def _params(tau, max_sifts, max_imf, min_extrema, end, B, S, EXACT, fp, threads=0):
    prm = _lib.default_params()
    prm.fp64 = 1 if fp in ("fp64", "float64", 64) else 0
    if fp not in ("fp64", "float64", 64, "fp32", "float32", 32):
        raise ValueError("fp must be 'fp64' or 'fp32'")
    if end not in _lib.END_MODES or end == "learned":
        raise ValueError("end must be 'natural' or 'mirror' (learned: see eemd4cuda_amd.learned)")
    prm.end_mode = _lib.END_MODES[end]
    prm.max_imf = int(max_imf)
    prm.max_sifts = int(max_sifts)
    prm.min_extrema = int(min_extrema)
    prm.block = 0 if EXACT else int(B)
    prm.halo = int(S)
    prm.tau = float(tau)
    prm.threads = int(threads)
    return prm


def gpu_eemd_full(x, E=64, tau=0.25, max_sifts=50, max_imf=12, seed=0, K=None, nthreads=128,
                  batch=None, B=250, S=30, EXACT=False, *, end="natural", min_extrema=4,
                  fp="fp64"):
    """Full EMD of every row of x (one ensemble member per row, noise already
    added).  Returns (imfs [E, maxm, N], nsift [E], nmode [E]) with maxm =
    max(nmode); IMF levels a member did not reach are zero."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        if E != 1:
            raise ValueError("1-D input x is a single trial (E=1). For E>1 pass a 2-D array.")
        x = x[None, :]
    if x.ndim != 2:
        raise ValueError("x must be [E, N]")
    E, N = x.shape
    if max_imf == -1:
        max_imf = 64
    prm = _params(tau, max_sifts, max_imf, min_extrema, end, B, S, EXACT, fp)
    # each row is its own signal with a single member and zero noise
    offsets = np.arange(E + 1, dtype=np.int64) * N
    out = _lib.engine().run_batch(x.reshape(-1), offsets, 1, np.zeros(E), prm,
                                  member_imfs=True)
    nmode = out["nmode"][:, 0].copy()
    nsift = out["nsift"][:, 0].copy()
    imfs = out["member_imfs"].reshape(E, max_imf, N).astype(np.float64, copy=False)
    maxm = int(nmode.max()) if nmode.size else 0
    return imfs[:, :maxm, :].copy(), nsift, nmode


def gpu_eemd(x, E=64, eps=0.2, tau=0.25, max_sifts=50, seed=0, K=None, nthreads=128, B=250,
             S=30, EXACT=False, *, end="natural", fp="fp64"):
    """First-IMF EEMD: numpy noise from default_rng(seed) scaled by eps*std(x),
    one IMF per member.  Returns (mean IMF-0, nsift [E], IMF-0 per member)."""
    x = np.asarray(x, dtype=np.float64)
    N = len(x)
    z = np.random.default_rng(seed).standard_normal((E, N))
    prm = _params(tau, max_sifts, 1, 4, end, B, S, EXACT, fp)
    out = _lib.engine().run_batch(x, np.array([0, N]), E, np.array([eps * np.std(x)]), prm,
                                  noise=z, member_imfs=True)
    imfs = out["member_imfs"].reshape(E, 1, N)[:, 0, :].astype(np.float64, copy=False)
    return imfs.mean(axis=0), out["nsift"][0].copy(), imfs


def gpu_emd_single(x, t=None, tau=0.25, max_sifts=50, K=None, nthreads=128, B=250, S=30,
                   EXACT=False, *, end="natural", fp="fp64"):
    """First IMF of one signal (no noise).  Returns (imf, n_sift)."""
    imfs, nsift, nmode = gpu_eemd_full(np.asarray(x, dtype=np.float64)[None, :], E=1, tau=tau,
                                       max_sifts=max_sifts, max_imf=1, B=B, S=S, EXACT=EXACT,
                                       end=end, fp=fp)
    N = len(x)
    imf = imfs[0, 0] if imfs.shape[1] else np.zeros(N)
    return imf, int(nsift[0])
