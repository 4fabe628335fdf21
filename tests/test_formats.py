"""imfs.npz, stft.npz, hilbert.npz and report.json (audio_eemd/report.py:55-215) and EMDResult npz round trips."""
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


def test_stft_and_hilbert_npz_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    f = np.arange(9) * 10.0
    t = np.arange(5) * 0.1
    z0 = rng.standard_normal((9, 5)) + 1j * rng.standard_normal((9, 5))
    zk = rng.standard_normal((3, 9, 5)) + 1j * rng.standard_normal((3, 9, 5))
    p = formats.write_stft_npz(tmp_path, (f, t, z0, zk))
    assert p.name == "stft.npz"
    f2, t2, z02, zk2 = formats.read_stft_npz(p)
    assert np.array_equal(f2, f) and np.array_equal(t2, t)
    assert np.array_equal(z02, z0) and np.array_equal(zk2, zk) and zk2.dtype == np.complex128
    with np.load(p, allow_pickle=False) as d:
        assert set(d.files) == {"freqs", "times", "Zxx_orig", "Zxx_imf"}
    hil = [(rng.standard_normal(50), rng.random(50)) for _ in range(4)]
    p = formats.write_hilbert_npz(tmp_path, hil)
    assert p.name == "hilbert.npz"
    back = formats.read_hilbert_npz(p)
    assert len(back) == 4
    for (w, a), (w2, a2) in zip(hil, back):
        assert np.array_equal(w, w2) and np.array_equal(a, a2)
    with np.load(p, allow_pickle=False) as d:
        assert sorted(d.files) == sorted(["inst_freq_%d" % k for k in range(4)] +
                                         ["inst_amp_%d" % k for k in range(4)])


def test_mode_table_closed_forms():
    sr, n = 48000, 4800                       # 10 Hz bins
    t = np.arange(n) / sr
    a1, a2 = 1.0, 0.5
    m1 = a1 * np.sin(2 * np.pi * 1000.0 * t + 0.3)   # 100 whole cycles, exact bin
    m2 = a2 * np.sin(2 * np.pi * 250.0 * t + 0.3)    # 25 whole cycles
    x = m1 + m2
    rows, rrow, q = formats.mode_table(np.stack([m1, m2]), np.zeros(n), x, sr)
    assert [r["k"] for r in rows] == [0, 1] and rrow["k"] == "residual"
    assert rows[0]["dominant_freq_hz"] == 1000.0 and rows[1]["dominant_freq_hz"] == 250.0
    # phase 0.3 keeps zeros between samples: 2 crossings per cycle
    assert rows[0]["n_zerocross"] == 200 and rows[1]["n_zerocross"] == 50
    # whole-cycle tones are orthogonal: energy fractions a_i^2 / sum a^2, summing to 1
    assert abs(rows[0]["energy_frac"] - a1 ** 2 / (a1 ** 2 + a2 ** 2)) < 1e-12
    assert abs(rows[1]["energy_frac"] - a2 ** 2 / (a1 ** 2 + a2 ** 2)) < 1e-12
    assert abs(rows[0]["std"] - a1 / np.sqrt(2)) < 1e-12
    assert rrow["energy_frac"] == 0.0 and rrow["dominant_freq_hz"] is None   # no energy
    assert rows[0]["sifting"] is None
    assert q == {"bit_depth": None, "floor_std": None, "n_flagged": 0}
    blk = formats.decomposition_block(2, 1e-15, rows, rrow, q)
    assert abs(blk["energy_sum"] - 1.0) < 1e-12 and blk["n_modes"] == 2
    # short series: no periodogram
    r_short, _, _ = formats.mode_table(np.ones((1, 10)), np.zeros(10), np.ones(10), sr)
    assert r_short[0]["dominant_freq_hz"] is None


def test_quantization_floor_and_flags():
    # 16-bit code at full scale 1: step 2^-15, uniform noise std step / sqrt(12)
    assert abs(formats.quantization_floor(1.0, 16) - 2.0 ** -15 / np.sqrt(12.0)) < 1e-20
    assert formats.quantization_floor(1.0, None) is None
    assert formats.quantization_floor(1.0, 32) is None
    n = 1000
    x = np.sin(np.arange(n) * 0.05)
    tiny = 1e-7 * np.sin(np.arange(n) * 0.9)       # below the 16-bit floor (~8.8e-6)
    rows, _, q = formats.mode_table(np.stack([x, tiny]), np.zeros(n), x, 48000, bit_depth=16)
    assert [r["quantization_level"] for r in rows] == [False, True]
    assert q["n_flagged"] == 1 and q["bit_depth"] == 16


def test_report_json(tmp_path):
    import json
    rows, rrow, q = formats.mode_table(np.ones((1, 64)) * np.sin(np.arange(64)), np.zeros(64),
                                       np.sin(np.arange(64)), 1000)
    rep = {"decomposition": formats.decomposition_block(1, 0.0, rows, rrow, q)}
    p = formats.write_report_json(tmp_path, rep)
    assert p.name == "report.json"
    assert json.loads(p.read_text()) == json.loads(json.dumps(rep))
