"""imfs.npz (audio_eemd/report.py:177-196) and EMDResult npz round trips."""
import numpy as np

from eemd4cuda_amd import formats
from eemd4cuda_amd.emd_par import EMDResult


def test_imfs_npz_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(500)
    imfs = rng.standard_normal((3, 500))
    resid = x - imfs.sum(axis=0)
    stats = dict(method="gpu", E=16, n_modes=3, batch=16, gpu_time_s=0.25, x=x,
                 per_trial_total_sifts=[30, 30, 28], max_imf_reached=False,
                 fp_mode="fp64", per_mode_sifts=np.arange(3))
    p = formats.write_imfs_npz(tmp_path, x, np.arange(500) / 48000.0, imfs, resid, 3, stats)
    assert p.name == "imfs.npz"
    d = formats.read_imfs_npz(p)
    assert set(d) == {"x", "t", "imfs", "residual", "n_modes", "stats"}
    assert np.array_equal(d["imfs"], imfs) and np.array_equal(d["residual"], resid)
    assert d["n_modes"] == 3
    # scalars and integer lists become stat_*; strings, arrays and the x copy do not
    assert d["stats"]["E"] == 16 and d["stats"]["max_imf_reached"] is False
    assert np.array_equal(d["stats"]["per_trial_total_sifts"], [30, 30, 28])
    assert "x" not in d["stats"] and "fp_mode" not in d["stats"] and "per_mode_sifts" not in d["stats"]
    with np.load(p, allow_pickle=False) as z:
        assert z["stat_per_trial_total_sifts"].dtype == np.int64
        assert z["n_modes"].dtype == np.int64


def test_emdresult_save_load(tmp_path):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(300)
    imfs = rng.standard_normal((2, 300))
    r = EMDResult(imfs=imfs, residual=x - imfs.sum(axis=0), n_modes=2,
                  stats=dict(method="gpu", E=4, x=x, gpu_time_s=0.1))
    assert r.recon_error() <= 1e-12
    path = tmp_path / "r.npz"
    r.save(str(path))
    q = EMDResult.load(str(path))
    assert q.n_modes == 2 and np.array_equal(q.imfs, imfs)
    assert np.array_equal(q.residual, r.residual)
