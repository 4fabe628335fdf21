"""CEEMDAN and first-IMF sifting on the MI355X engine.

Mirrors emd_ref.ceemdan (emd_ref.py:220-259) and emd_ref.emd_1imf
(emd_ref.py:158-162), including the reference's exact semantics:

* stage 0 decomposes each of the E noise realisations W[e]
  (default_rng(seed).standard_normal((E, N)), emd_ref.py:167-169).  Only
  their IMF counts are used, to decide when to stop;
* stage k averages, over e, the first IMF of resid + eps*std(resid)*W[e].
  Like the reference, it uses the raw noise W[e] at every stage;
* the loop stops when resid is ~0 (max|resid| < 1e-12 or std < 1e-10), or
  after the stage whose index reaches every noise's IMF count.

Each stage is one batched engine call (E members, max_imf=1, explicit noise
matrix W), and stage 0 is one batched call over the E noise signals.  The
stage loop itself is inherently sequential in the IMF index.
"""
from __future__ import annotations

import numpy as np

from . import _lib

__all__ = ["ceemdan", "emd_1imf"]


def _params(fp, tau, max_sifts, max_imf, end, B, S, min_extrema=4):
    prm = _lib.default_params()
    prm.fp64 = 1 if fp == "fp64" else 0
    prm.end_mode = _lib.END_MODES[end]
    prm.max_imf = int(max_imf)
    prm.max_sifts = int(max_sifts)
    prm.min_extrema = int(min_extrema)
    prm.tau = float(tau)
    prm.block = int(B)
    prm.halo = int(S) if fp == "fp64" or S != 30 else 14
    prm.members_per_task = 1
    return prm


def emd_1imf(x, tau=0.25, max_sifts=50, end="natural", fp="fp64", B=1024, S=30, engine=None):
    """First IMF only (emd_ref.py:158-162): returns (imf, n_sift)."""
    eng = engine or _lib.engine()
    x = np.asarray(x, dtype=np.float64)
    N = len(x)
    out = eng.run_batch(x, np.array([0, N]), 1, np.zeros(1),
                        _params(fp, tau, max_sifts, 1, end, B, S), member_imfs=True)
    if out["nmode"][0, 0] < 1:
        raise ValueError("the signal has fewer than 4 extrema: no IMF")
    imf = np.asarray(out["member_imfs"][:N], np.float64)
    return imf, int(out["nsift"][0, 0])


def ceemdan(x, E=100, eps=0.2, seed=0, tau=0.25, max_sifts=50, end="natural", fp="fp64",
            B=1024, S=30, noise=None, engine=None):
    """CEEMDAN of x (emd_ref.py:220-259).

    Returns (imfs [K, N], resid, info): info holds E, n_imfs, the noise IMF
    counts and the per-stage, per-member first-IMF sift counts.  `noise`
    overrides the default_rng(seed) noise matrix [E, N].  B / S: decay-seamed
    block-Thomas geometry (B <= 0: exact Thomas, the reference's solver).
    """
    eng = engine or _lib.engine()
    x = np.asarray(x, dtype=np.float64)
    if x.ndim != 1 or len(x) < 8:
        raise ValueError("x must be 1-D with at least 8 samples")
    N = len(x)
    W = (np.random.default_rng(seed).standard_normal((E, N)) if noise is None
         else np.ascontiguousarray(noise, dtype=np.float64))
    if W.shape != (E, N):
        raise ValueError("noise must be [E, N]")
    # stage 0: the E noise realisations as E separate signals, plain EMD
    offs = np.arange(E + 1, dtype=np.int64) * N
    nz = eng.run_batch(W.ravel(), offs, 1, np.zeros(E), _params(fp, tau, max_sifts, 64, end, B, S))
    noise_n = nz["nmode"][:, 0].astype(np.int64)
    prm1 = _params(fp, tau, max_sifts, 1, end, B, S)
    resid = x.copy()
    imfs, stage_sifts = [], []
    while True:
        if np.max(np.abs(resid)) < 1e-12 or np.std(resid) < 1e-10:
            break
        stage = len(imfs)
        beta = eps * np.std(resid)
        out = eng.run_batch(resid, np.array([0, N]), E, np.array([beta]), prm1, noise=W)
        if np.any(out["nmode"][0] < 1):
            raise RuntimeError("a CEEMDAN trial produced no IMF")
        c = out["avg"][:N].copy()
        imfs.append(c)
        stage_sifts.append(out["nsift"][0].copy())
        resid = resid - c
        if np.all(stage >= noise_n):
            break
    info = dict(E=E, n_imfs=len(imfs), noise_nimf=noise_n.astype(np.int32),
                stage_sifts=np.array(stage_sifts, np.int32))
    return np.array(imfs), resid, info
