"""numpy restatement of the reference's CPU EEMD path -- TEST INFRASTRUCTURE ONLY.

Used as (1) the timed CPU baseline of ``bench.py`` (the reference's own CPU
path, restated so it can run where the reference is absent) and (2) a second
oracle pinned against golden vectors written by the reference itself.  The
shipped package never imports it.

It uses the same numpy operations, in the same order, as ``emd_ref.py`` so that
its rounding (numpy's SIMD ``power`` in the cube, pairwise ``np.sum``,
``searchsorted``) is the reference's:

* ``_system`` / ``_rhs``   <- emd_ref.py:31-48
* ``_thomas``              <- emd_ref.py:50-65 (a scalar Python loop, as there)
* ``_eval``                <- emd_ref.py:67-75
* ``envelope``             <- emd_ref.py:77-103
* ``emd_1d``               <- emd_ref.py:108-156
* ``eemd_member`` / ``eemd`` <- emd_ref.py:171-177, 186-218 (max_imf honoured
  per member; ProcessPoolExecutor with as_completed scheduling)
"""
from __future__ import annotations

import os
import time

import numpy as np


def _system(T):
    H = np.diff(T)
    n = len(T) - 2
    d = 2.0 * (H[0:n] + H[1:n + 1])
    l = np.zeros(n)
    u = np.zeros(n)
    l[1:] = H[1:n]
    u[0:n - 1] = H[1:n]
    return d, l, u, H, n


def _rhs(Y, H, n):
    return 6.0 * ((Y[2:n + 2] - Y[1:n + 1]) / H[1:n + 1] - (Y[1:n + 1] - Y[0:n]) / H[0:n])


def _thomas(d, l, u, r):
    n = len(d)
    if n == 0:
        return np.zeros(0)
    cp = np.empty(n)
    dp = np.empty(n)
    cp[0] = (u[0] if n > 1 else 0.0) / d[0]
    dp[0] = r[0] / d[0]
    for i in range(1, n):
        den = d[i] - l[i] * cp[i - 1]
        cp[i] = (u[i] if i < n - 1 else 0.0) / den
        dp[i] = (r[i] - l[i] * dp[i - 1]) / den
    out = np.empty(n)
    out[-1] = dp[-1]
    for i in range(n - 2, -1, -1):
        out[i] = dp[i] - cp[i] * out[i + 1]
    return out


def _eval(T, Y, M, t):
    P = len(T)
    j = np.clip(np.searchsorted(T, t) - 1, 0, P - 2)
    h = T[j + 1] - T[j]
    a = (T[j + 1] - t) / h
    b = (t - T[j]) / h
    return (a * Y[j] + b * Y[j + 1]
            + ((a ** 3 - a) * M[j] + (b ** 3 - b) * M[j + 1]) * (h * h) / 6.0)


def _knots(x, want_max):
    if want_max:
        return np.where((x[1:-1] > x[:-2]) & (x[1:-1] > x[2:]))[0] + 1
    return np.where((x[1:-1] < x[:-2]) & (x[1:-1] < x[2:]))[0] + 1


def envelope(t, x, want_max, end="natural"):
    k = _knots(x, want_max)
    p = len(k)
    if p < 2:
        return np.full_like(t, x[k[0]] if p == 1 else 0.0), p
    T, Y = t[k], x[k]
    if end == "mirror" and p >= 3:
        T = np.concatenate(([T[0] - (T[1] - T[0])], T, [T[-1] + (T[-1] - T[-2])]))
        Y = np.concatenate(([Y[1]], Y, [Y[-2]]))
    d, l, u, H, n = _system(T)
    M = np.zeros(len(T))
    M[1:-1] = _thomas(d, l, u, _rhs(Y, H, n))
    return _eval(T, Y, M, t), p


def _count(x, want_max):
    if want_max:
        return int(np.sum((x[1:-1] > x[:-2]) & (x[1:-1] > x[2:])))
    return int(np.sum((x[1:-1] < x[:-2]) & (x[1:-1] < x[2:])))


def emd_1d(x, tau=0.25, max_sifts=50, max_imf=-1, min_extrema=4, end="natural"):
    x = np.asarray(x, dtype=float)
    t = np.arange(len(x), dtype=float)
    resid = x.copy()
    imfs, per_mode = [], []
    while True:
        if _count(resid, True) + _count(resid, False) < min_extrema:
            break
        if max_imf != -1 and len(imfs) >= max_imf:
            break
        h = resid.copy()
        prev = None
        ns = 0
        for _ in range(max_sifts):
            if _count(h, True) < 1 or _count(h, False) < 1:
                break
            eu, _ = envelope(t, h, True, end)
            el, _ = envelope(t, h, False, end)
            hn = h - 0.5 * (eu + el)
            ns += 1
            if prev is not None:
                den = np.sum(prev ** 2)
                sd = np.sum((prev - hn) ** 2) / den if den > 0 else 0.0
                if sd < tau:
                    h = hn
                    break
            prev = hn
            h = hn
        imfs.append(h)
        per_mode.append(ns)
        resid = resid - h
    return imfs, resid, dict(n_modes=len(imfs), total_sifts=int(sum(per_mode)),
                             per_mode_sifts=per_mode)


def eemd_member(args):
    """One ensemble member: EMD of x + z*sigma (emd_ref.py:171-177)."""
    x, z, sigma, tau, max_sifts, max_imf, min_extrema, end = args
    t0 = time.perf_counter()
    imfs, resid, st = emd_1d(x + z * sigma, tau=tau, max_sifts=max_sifts, max_imf=max_imf,
                             min_extrema=min_extrema, end=end)
    return imfs, st, time.perf_counter() - t0


def eemd(x, z, sigma, tau=0.0, max_sifts=10, max_imf=8, min_extrema=4, end="mirror",
         workers=None):
    """EEMD over the rows of z with level-wise sequential averaging.

    workers: None/0 -> serial; k>0 -> ProcessPoolExecutor(k) with as_completed
    dynamic scheduling (emd_ref.py:186-197)."""
    x = np.asarray(x, dtype=float)
    E, N = z.shape
    tasks = [(x, z[e], sigma, tau, max_sifts, max_imf, min_extrema, end) for e in range(E)]
    if not workers:
        res = [eemd_member(a) for a in tasks]
    else:
        from concurrent.futures import ProcessPoolExecutor, as_completed
        res = [None] * E
        with ProcessPoolExecutor(max_workers=workers) as ex:
            futs = {ex.submit(eemd_member, a): i for i, a in enumerate(tasks)}
            for f in as_completed(futs):
                res[futs[f]] = f.result()
    maxm = max((len(r[0]) for r in res), default=0)
    avg = np.zeros((maxm, N))
    cnt = np.zeros(maxm)
    for r in res:
        for i, imf in enumerate(r[0]):
            avg[i] += imf
            cnt[i] += 1
    avg = avg / np.where(cnt > 0, cnt, 1)[:, None]
    resid = x - avg.sum(axis=0)
    return dict(avg=avg, resid=resid, n_modes=maxm,
                nmode=np.array([len(r[0]) for r in res], dtype=np.int32),
                nsift=np.array([r[1]["total_sifts"] for r in res], dtype=np.int32),
                per_mode=[r[1]["per_mode_sifts"] for r in res],
                member_s=np.array([r[2] for r in res]))


def cpu_count() -> int:
    return os.cpu_count() or 1
