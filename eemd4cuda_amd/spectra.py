"""Time-frequency views of IMFs on the MI355X engine (audio_eemd/spectra.py:18-63).

stft_block(x, sr, nperseg, noverlap) returns (f, t, Zxx) like the reference's
scipy.signal.stft call (eemd_stft: windowed frames, batched hipFFT, scaling).

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

__all__ = ["hilbert_if", "if_stats", "stft_block"]


def _bind():
    L = _lib.load()
    if not hasattr(L, "_spectra_bound"):
        L.eemd_hilbert_if.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_double,
                                      C.c_void_p, C.c_void_p]
        L.eemd_hilbert_if.restype = C.c_int
        L.eemd_stft.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int32,
                                C.c_int32, C.c_void_p]
        L.eemd_stft.restype = C.c_int
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


def stft_block(x, sr, nperseg: int = 4096, noverlap: int | None = None, engine=None):
    """STFT of a 1-D signal, or of each row of a 2-D block (audio_eemd/spectra.py:18-29):
    (f, t, Zxx) as scipy.signal.stft(x, fs=sr, window='hann', nperseg, noverlap,
    boundary='zeros', padded=False) returns them."""
    eng = engine or _lib.engine()
    L = _bind()
    if noverlap is None:
        noverlap = nperseg // 2
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.ndim not in (1, 2):
        raise ValueError("x must be 1-D or 2-D")
    rows = a.reshape(-1, a.shape[-1])
    n = rows.shape[1]
    step = nperseg - noverlap
    if nperseg < 2 or not 0 <= noverlap < nperseg:
        raise ValueError("need nperseg >= 2 and 0 <= noverlap < nperseg")
    if n + 2 * (nperseg // 2) < nperseg:
        raise ValueError("signal shorter than one segment")
    nseg = (n + 2 * (nperseg // 2) - noverlap) // step
    nf = nperseg // 2 + 1
    z = np.empty((rows.shape[0], nf, nseg), dtype=np.complex128)
    rc = L.eemd_stft(eng.h, rows.ctypes.data_as(C.c_void_p), rows.shape[0], n, int(nperseg),
                     int(noverlap), z.ctypes.data_as(C.c_void_p))
    eng._check(rc, "eemd_stft")
    f = np.arange(nf) * (float(sr) / nperseg)
    t = np.arange(nseg) * step / float(sr)
    return f, t, (z[0] if a.ndim == 1 else z)


def if_stats(w: np.ndarray) -> dict:
    """Median / IQR of an IF series, positive part only (audio_eemd/spectra.py:51-63)."""
    wp = w[w > 0]
    if wp.size == 0:
        return {"median": None, "iqr": None, "n_pos_frac": 0.0}
    lo, hi = np.percentile(wp, [25, 75])
    return {"median": float(np.median(wp)), "iqr": float(hi - lo),
            "n_pos_frac": float(wp.size / w.size)}
