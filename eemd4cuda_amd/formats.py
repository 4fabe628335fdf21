"""On-disk formats of the decomposition (the reference's artifact layout).

* write_imfs_npz / read_imfs_npz (audio_eemd/report.py:177-196): one
  compressed npz with x, t, imfs [n_modes, N], residual, n_modes (int64) and
  every scalar or integer-list entry of the stats dict as stat_<key> (the
  input copy stats['x'] is not repeated).
* write_stft_npz / write_hilbert_npz and their readers
  (audio_eemd/report.py:198-215): the spectra of the IMFs (freqs, times,
  Zxx_orig [F, T], Zxx_imf [k, F, T]; inst_freq_k / inst_amp_k per IMF).
* mode_table / decomposition_block / write_report_json
  (audio_eemd/report.py:55-135, 161-164): the per-IMF table and the
  "decomposition" section of report.json.  The report's file metadata,
  audio statistics and caveats belong to the application around the
  decomposition and are the caller's.

EMDResult.save / load (emd_par.py:84-97) live in emd_par.  Only inputs and
outputs are stored: no pickles.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

__all__ = ["write_imfs_npz", "read_imfs_npz", "write_stft_npz", "read_stft_npz",
           "write_hilbert_npz", "read_hilbert_npz", "mode_table", "quantization_floor",
           "decomposition_block", "write_report_json"]


def write_imfs_npz(outdir, x, t, imfs, residual, n_modes, stats) -> Path:
    """imfs.npz: x, t, imfs [n_modes, N], residual, n_modes, scalar stats."""
    payload = {"x": np.asarray(x), "t": np.asarray(t), "imfs": np.asarray(imfs),
               "residual": np.asarray(residual), "n_modes": np.int64(n_modes)}
    for k, v in (stats or {}).items():
        if k == "x":
            continue
        if isinstance(v, (int, float, bool, np.integer, np.floating, np.bool_)):
            payload["stat_" + k] = np.asarray(v)
        elif isinstance(v, (list, tuple)) and all(isinstance(i, (int, np.integer)) for i in v):
            payload["stat_" + k] = np.asarray(v, dtype=np.int64)
    outdir = Path(outdir)
    outdir.mkdir(parents=True, exist_ok=True)
    p = outdir / "imfs.npz"
    np.savez_compressed(p, **payload)
    return p


def read_imfs_npz(path) -> dict:
    """The arrays of an imfs.npz plus a 'stats' dict rebuilt from stat_* keys."""
    with np.load(path, allow_pickle=False) as d:
        out = {k: d[k] for k in d.files if not k.startswith("stat_")}
        out["stats"] = {k[5:]: (d[k].item() if d[k].ndim == 0 else d[k]) for k in d.files
                        if k.startswith("stat_")}
    out["n_modes"] = int(out["n_modes"])
    return out


def _outdir(outdir) -> Path:
    d = Path(outdir)
    d.mkdir(parents=True, exist_ok=True)
    return d


def write_stft_npz(outdir, stft) -> Path:
    """stft.npz from (f, t, Zxx_orig [F, T], Zxx_imf [k, F, T])."""
    f, t, z_orig, z_imf = stft
    p = _outdir(outdir) / "stft.npz"
    np.savez_compressed(p, freqs=np.asarray(f), times=np.asarray(t), Zxx_orig=np.asarray(z_orig),
                        Zxx_imf=np.asarray(z_imf))
    return p


def read_stft_npz(path):
    """(f, t, Zxx_orig, Zxx_imf) of a stft.npz."""
    with np.load(path, allow_pickle=False) as d:
        return d["freqs"], d["times"], d["Zxx_orig"], d["Zxx_imf"]


def write_hilbert_npz(outdir, hilb) -> Path:
    """hilbert.npz from a list of (inst_freq [N], inst_amp [N]), one per IMF."""
    payload = {}
    for k, (w, amp) in enumerate(hilb):
        payload["inst_freq_%d" % k] = np.asarray(w)
        payload["inst_amp_%d" % k] = np.asarray(amp)
    p = _outdir(outdir) / "hilbert.npz"
    np.savez_compressed(p, **payload)
    return p


def read_hilbert_npz(path):
    """The list of (inst_freq, inst_amp) of a hilbert.npz, in IMF order."""
    with np.load(path, allow_pickle=False) as d:
        k = len([n for n in d.files if n.startswith("inst_freq_")])
        return [(d["inst_freq_%d" % i], d["inst_amp_%d" % i]) for i in range(k)]


def _dominant_freq(x, sr):
    """Peak of the Hann-windowed periodogram, DC excluded; None when the
    series is too short or has no energy (audio_eemd/report.py:55-63)."""
    if len(x) < 16:
        return None
    spec = np.abs(np.fft.rfft(x * np.hanning(len(x))))
    if spec.sum() <= 0:
        return None
    return float(np.fft.rfftfreq(len(x), 1.0 / sr)[1 + int(np.argmax(spec[1:]))])


def _zero_crossings(x) -> int:
    """Sign-bit changes between consecutive samples (audio_eemd/report.py:66-67)."""
    return int(np.count_nonzero(np.diff(np.signbit(x))))


def quantization_floor(peak, bit_depth):
    """Std of uniform quantisation noise of a bit_depth-bit code with full
    scale `peak`: step / (2 sqrt 3), step = peak / 2^(b-1); None for float
    formats (no bit depth, or 32 bits and more) (audio_eemd/report.py:70-78)."""
    if not bit_depth or bit_depth >= 32:
        return None
    return float(peak / 2.0 ** (bit_depth - 1) / (2.0 * np.sqrt(3.0)))


def mode_table(imfs, residual, x, sr, bit_depth=None):
    """Per-IMF statistics rows, the residual row and the quantisation summary
    (audio_eemd/report.py:81-110): k, std, energy_frac = ||m||^2 / ||x||^2,
    dominant_freq_hz, n_zerocross, sifting (None: not per mode) and
    quantization_level (std below the bit depth's noise floor)."""
    x = np.asarray(x, dtype=np.float64)
    e_x = float(np.sum(x ** 2))
    floor = quantization_floor(float(np.abs(x).max()), bit_depth)

    def row(k, m):
        m = np.asarray(m, dtype=np.float64)
        sd = float(np.std(m))
        return {"k": k, "std": sd,
                "energy_frac": float(np.sum(m ** 2)) / e_x if e_x > 0 else 0.0,
                "dominant_freq_hz": _dominant_freq(m, sr),
                "n_zerocross": _zero_crossings(m),
                "sifting": None,
                "quantization_level": bool(floor is not None and sd < floor)}

    rows = [row(k, m) for k, m in enumerate(np.asarray(imfs))]
    rrow = row("residual", residual)
    qinfo = {"bit_depth": bit_depth, "floor_std": floor,
             "n_flagged": int(sum(r["quantization_level"] for r in rows))}
    return rows, rrow, qinfo


def decomposition_block(n_modes, recon_error, rows, rrow, qinfo) -> dict:
    """The "decomposition" section of report.json (audio_eemd/report.py:125-133);
    energy_sum = sum of the IMF and residual energy fractions (not 1 in
    general: EMD modes are not orthogonal)."""
    return {"n_modes": int(n_modes), "recon_error": float(recon_error), "imf": rows,
            "residual": rrow,
            "energy_sum": sum(r["energy_frac"] for r in rows) + rrow["energy_frac"],
            "quantization": qinfo}


def write_report_json(outdir, report: dict) -> Path:
    """report.json, indented like the reference's (audio_eemd/report.py:161-164)."""
    p = _outdir(outdir) / "report.json"
    p.write_text(json.dumps(report, indent=2), encoding="utf-8")
    return p
