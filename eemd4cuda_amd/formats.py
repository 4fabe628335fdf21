"""On-disk formats of the decomposition (the reference's artifact layout).

write_imfs_npz / read_imfs_npz follow audio_eemd/report.py:177-196: one
compressed npz with x, t, imfs [n_modes, N], residual, n_modes (int64) and
every scalar or integer-list entry of the stats dict as stat_<key> (the input
copy stats['x'] is not repeated).  EMDResult.save / load (emd_par.py:84-97)
live in emd_par.  Only inputs and outputs are stored: no pickles.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

__all__ = ["write_imfs_npz", "read_imfs_npz"]


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
