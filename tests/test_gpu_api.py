"""The reference's call surface on the MI355X engine: run_eemd / decompose /
gpu_eemd_full, mirroring audio_eemd/tests/test_eemd_stage.py and the A1-A5
checks of test_emd_gpu.py (GPU backend instead of CPU)."""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.fixture
def two_tones():
    sr = 8000
    t = np.arange(4 * sr) / sr
    return (np.sin(2 * np.pi * 100 * t) + np.sin(2 * np.pi * 105 * t)) / 2, sr


def test_emd_single(two_tones):
    from eemd4cuda_amd.eemd_stage import run_eemd
    x, sr = two_tones
    d = run_eemd(x, method="gpu", E=1, seed=0)
    assert d.n_modes >= 1
    assert d.imfs.shape == (d.n_modes, len(x))
    assert d.residual.shape == (len(x),)
    assert d.recon < 1e-6 * np.abs(x).max()
    assert d.stats["method"] == "gpu" and d.stats["E"] == 1


def test_eemd_ensemble(two_tones):
    from eemd4cuda_amd.eemd_stage import run_eemd
    x, sr = two_tones
    d = run_eemd(x, method="gpu", E=8, seed=1)
    assert d.n_modes >= 1 and d.imfs.shape == (d.n_modes, len(x))
    assert d.recon < 1e-6 * np.abs(x).max()
    assert len(d.stats["per_trial_total_sifts"]) == 8


def test_seed_determinism(two_tones):
    from eemd4cuda_amd.eemd_stage import run_eemd
    x, sr = two_tones
    a = run_eemd(x, method="gpu", E=4, seed=42)
    b = run_eemd(x, method="gpu", E=4, seed=42)
    assert a.n_modes == b.n_modes
    assert np.array_equal(a.imfs, b.imfs) and np.array_equal(a.residual, b.residual)


def test_decompose_matches_reference_host_driver(two_tones):
    """decompose(method='gpu') = the reference's _gpu_decompose: numpy noise
    default_rng(seed).standard_normal((E,N)) * eps*std(x), per-member EMD,
    level-wise mean (emd_par.py:225-250); here checked against the oracle."""
    from eemd4cuda_amd.emd_par import decompose
    x, _ = two_tones
    x = x[:6000]
    r = decompose(x, method="gpu", E=4, eps=0.2, tau=0.0, max_sifts=10, max_imf=6, seed=3,
                  end="mirror", B=0)
    z = np.random.default_rng(3).standard_normal((4, len(x)))
    o = orc.eemd(x, z, 0.2 * np.std(x), tau=0.0, max_sifts=10, max_imf=6, end="mirror")
    assert r.n_modes == o["n_modes"]
    assert np.max(np.abs(r.imfs - o["avg"])) <= 1e-12
    assert r.recon_error() <= 1e-12
    st = r.stats
    for k in ("method", "E", "n_modes", "batch", "per_trial_total_sifts", "gpu_time_s", "x",
              "total_sifts", "max_imf_reached", "fp_mode", "boundary", "B", "S", "wg",
              "members_per_wg", "hbm_bytes", "sample_sifts"):
        assert k in st, k
    assert st["sample_sifts"] == int(np.sum(o["nsift"])) * len(x)
    assert st["hbm_bytes"] == int(st["sample_sifts"] * 19.2)
    assert 1 <= st["members_per_wg"] <= 4
    assert np.array_equal(st["per_trial_per_mode_sifts"], o["per_mode"][:, :r.n_modes])


def test_gpu_eemd_full_single_sift():
    """A1 (test_emd_gpu.py:89-102): one sift of the chirp vs the oracle's envelopes."""
    from eemd4cuda_amd import signals
    from eemd4cuda_amd.emd_gpu import gpu_emd_single
    xs = signals.chirp2048()
    imf, ns = gpu_emd_single(xs, tau=0.25, max_sifts=1, EXACT=True)
    eu, _ = orc.envelope(xs, True, "natural")
    el, _ = orc.envelope(xs, False, "natural")
    assert ns == 1
    assert np.max(np.abs(imf - (xs - 0.5 * (eu + el)))) <= 1e-13


def test_gpu_eemd_full_shapes_and_guard():
    from eemd4cuda_amd.emd_gpu import gpu_eemd_full
    rng = np.random.default_rng(7)
    x = np.sin(np.arange(3000) * 0.05)[None, :] + 0.05 * rng.standard_normal((4, 3000))
    imfs, nsift, nmode = gpu_eemd_full(x, E=4, tau=0.25, max_imf=8)
    assert imfs.shape == (4, nmode.max(), 3000)
    for e in range(4):
        ref, _, st = orc.emd_1d(x[e], tau=0.25, max_sifts=50, max_imf=8, end="natural", B=250, S=30)
        assert nmode[e] == st["n_modes"]
        assert np.max(np.abs(imfs[e, :nmode[e]] - ref)) <= 1e-12
    with pytest.raises(ValueError):
        gpu_eemd_full(x[0], E=4)


def test_save_load_roundtrip(tmp_path, two_tones):
    from eemd4cuda_amd.emd_par import EMDResult, decompose
    x, _ = two_tones
    r = decompose(x[:4000], method="gpu", E=2, seed=0)
    p = tmp_path / "r.npz"
    r.save(str(p))
    q = EMDResult.load(str(p))
    assert np.array_equal(q.imfs, r.imfs) and q.n_modes == r.n_modes


@pytest.mark.parametrize("noise", ["numpy", "philox"])
def test_decompose_batches_equal_one_call(noise):
    """Members split over several engine calls (emd_par.py:234-250 batching)
    give the same members and, up to re-association of the per-level sums,
    the same ensemble mean as one call."""
    from eemd4cuda_amd import emd_par
    x = np.sin(np.arange(3000) * 0.05) + 0.3 * np.sin(np.arange(3000) * 0.71)
    print("case: decompose batches noise=%s" % noise, flush=True)
    kw = dict(method="gpu", E=10, eps=0.2, tau=0.0, max_sifts=10, max_imf=6, end="mirror",
              seed=3, noise=noise, B=0)
    one = emd_par.decompose(x, batch=10, **kw)
    many = emd_par.decompose(x, batch=3, **kw)
    assert one.stats["n_batches"] == 1 and many.stats["n_batches"] == 4
    assert np.array_equal(one.stats["per_trial_total_sifts"], many.stats["per_trial_total_sifts"])
    assert one.n_modes == many.n_modes
    assert np.max(np.abs(one.imfs - many.imfs)) <= 1e-13 * np.max(np.abs(x))
    assert many.stats["max_imf_reached"] is False
    assert many.recon_error() <= 1e-12


def test_decompose_auto_batch_uses_free_memory():
    from eemd4cuda_amd import _lib, emd_par
    free, total = _lib.engine().mem_info()
    assert 0 < free <= total
    x = np.sin(np.arange(2000) * 0.05)
    r = emd_par.decompose(x, method="gpu", E=4, max_imf=-1, tau=0.0, max_sifts=10, end="mirror")
    assert r.stats["batch"] == 4        # four members fit at once
    assert r.stats["max_imf_reached"] is False


def test_integration_stub_runs_and_matches_oracle():
    """The ctypes stub of INTEGRATION.md, executed as written (library path
    filled in): its EEMD equals the oracle with the reference's DSBT geometry
    B=250, S=30."""
    import re
    from pathlib import Path

    from eemd4cuda_amd.build import lib_path
    txt = (Path(__file__).resolve().parents[1] / "INTEGRATION.md").read_text()
    code = re.search(r"```python\n(import ctypes as C.*?)```", txt, flags=re.S).group(1)
    code = code.replace('C.CDLL("libeemd_mi355x.so")', "C.CDLL(%r)" % str(lib_path()))
    ns = {}
    exec(compile(code, "INTEGRATION.md", "exec"), ns)
    rng = np.random.default_rng(11)
    N, E = 5000, 4
    x = np.sin(np.arange(N) * 0.02) + 0.4 * np.sin(np.arange(N) * 0.31) + 0.1 * rng.standard_normal(N)
    noise = np.random.default_rng(0).standard_normal((E, N))
    sigma = 0.2 * np.std(x)
    print("case: INTEGRATION.md stub", flush=True)
    imfs, resid, k, nsift, nmode = ns["_mi355x_eemd"](x, noise, sigma, 0.25, 50, 6, end="natural")
    o = orc.eemd(x, noise, sigma, tau=0.25, max_sifts=50, max_imf=6, end="natural", B=250, S=30)
    assert k == o["n_modes"]
    assert np.array_equal(nsift, o["nsift"]) and np.array_equal(nmode, o["nmode"])
    assert np.max(np.abs(imfs - o["avg"])) <= 1e-12
    assert np.max(np.abs(resid - o["resid"])) <= 1e-12
