"""Hilbert spectrum of IMFs on the MI355X engine (audio_eemd/spectra.py:32-63).

hilbert_if(imf, sr) returns (inst_freq, inst_amp) like the reference: z =
imf + i H(imf) by the FFT method of scipy.signal.hilbert, inst_freq =
np.gradient(np.unwrap(np.angle(z)), 1/sr) / 2 pi, inst_amp = |z|.  A 2-D
input is a batch of rows (for example all IMFs of a decomposition), done in
one call (eemd_hilbert_if: hipFFT for the transforms, HIP kernels for the
one-sided spectrum, phase, unwrapping scan and gradient).  if_stats is the
reference's host-side summary of an IF series.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

__all__ = ["hilbert_if", "if_stats"]


def _bind():
    L = _lib.load()
    if not hasattr(L, "_spectra_bound"):
        L.eemd_hilbert_if.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_double,
                                      C.c_void_p, C.c_void_p]
        L.eemd_hilbert_if.restype = C.c_int
        L._spectra_bound = True
    return L


def hilbert_if(imf, sr, engine=None):
    """Instantaneous frequency [Hz] and amplitude of one IMF ([N]) or of each
    row of a [rows, N] block."""
    eng = engine or _lib.engine()
    L = _bind()
    a = np.ascontiguousarray(imf, dtype=np.float64)
    if a.ndim not in (1, 2):
        raise ValueError("imf must be 1-D or 2-D")
    rows = a.reshape(-1, a.shape[-1])
    if rows.shape[1] < 2:
        raise ValueError("need at least 2 samples")
    f = np.empty_like(rows)
    amp = np.empty_like(rows)
    rc = L.eemd_hilbert_if(eng.h, rows.ctypes.data_as(C.c_void_p), rows.shape[0], rows.shape[1],
                           float(sr), f.ctypes.data_as(C.c_void_p), amp.ctypes.data_as(C.c_void_p))
    eng._check(rc, "eemd_hilbert_if")
    if a.ndim == 1:
        return f[0], amp[0]
    return f, amp


def if_stats(w: np.ndarray) -> dict:
    """Median / IQR of an IF series, positive part only (audio_eemd/spectra.py:51-63)."""
    wp = w[w > 0]
    if wp.size == 0:
        return {"median": None, "iqr": None, "n_pos_frac": 0.0}
    lo, hi = np.percentile(wp, [25, 75])
    return {"median": float(np.median(wp)), "iqr": float(hi - lo),
            "n_pos_frac": float(wp.size / w.size)}
