"""CEEMDAN (emd_ref.py:220-259) and first-IMF sifting (emd_ref.py:158-162).

The oracle restatement (oracle.orc.ceemdan) is pinned to the reference's own
output: tests/golden/golden_ceemdan.npz holds emd_ref.ceemdan on a 1024-sample
chirp, E=4, seed 0, with natural ends (tau 0.25, 50 sifts) and mirror ends
(tau 0.2, 20 sifts), plus every stage's per-trial sift counts and the noise
IMF counts.  The GPU path must reproduce both.
"""
import numpy as np
import pytest

from oracle import orc

CASES = [("natural", 0.25, 50), ("mirror", 0.2, 20)]


def _noise(g):
    return np.random.default_rng(int(g["seed"])).standard_normal((int(g["E"]), len(g["x"])))


@pytest.mark.parametrize("end,tau,max_sifts", CASES)
def test_oracle_ceemdan_matches_reference(golden, end, tau, max_sifts):
    g = golden("golden_ceemdan.npz")
    imfs, resid, info = orc.ceemdan(g["x"], _noise(g), eps=float(g["eps"]), tau=tau,
                                    max_sifts=max_sifts, end=end)
    k = end + "_"
    assert np.array_equal(info["noise_nimf"], g[k + "noise_nimf"])
    assert np.array_equal(info["stage_sifts"], g[k + "stage_sifts"])
    assert imfs.shape == g[k + "imfs"].shape
    scale = np.max(np.abs(g["x"]))
    assert np.max(np.abs(imfs - g[k + "imfs"])) <= 1e-10 * scale
    assert np.max(np.abs(resid - g[k + "resid"])) <= 1e-10 * scale
    # partition of unity of the staircase
    assert np.max(np.abs(imfs.sum(axis=0) + resid - g["x"])) <= 1e-12 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("end,tau,max_sifts", CASES)
def test_gpu_ceemdan_matches_reference_and_oracle(golden, end, tau, max_sifts):
    from eemd4cuda_amd import ceemdan
    g = golden("golden_ceemdan.npz")
    print("case: ceemdan N=1024 E=4 %s" % end, flush=True)
    imfs, resid, info = ceemdan.ceemdan(g["x"], E=int(g["E"]), eps=float(g["eps"]),
                                        seed=int(g["seed"]), tau=tau, max_sifts=max_sifts,
                                        end=end, B=0)
    k = end + "_"
    assert np.array_equal(info["noise_nimf"], g[k + "noise_nimf"])
    assert np.array_equal(info["stage_sifts"], g[k + "stage_sifts"])
    scale = np.max(np.abs(g["x"]))
    assert np.max(np.abs(imfs - g[k + "imfs"])) <= 1e-10 * scale
    assert np.max(np.abs(resid - g[k + "resid"])) <= 1e-10 * scale
    o_imfs, o_resid, _ = orc.ceemdan(g["x"], _noise(g), eps=float(g["eps"]), tau=tau,
                                     max_sifts=max_sifts, end=end)
    assert np.max(np.abs(imfs - o_imfs)) <= 1e-13 * scale
    assert np.max(np.abs(resid - o_resid)) <= 1e-13 * scale


@pytest.mark.gpu
def test_gpu_first_imf_matches_oracle():
    from eemd4cuda_amd import ceemdan, signals
    x = signals.chirp2048()
    print("case: emd_1imf N=2048", flush=True)
    imf, ns = ceemdan.emd_1imf(x, tau=0.25, max_sifts=50, end="natural", B=0)
    o, _, st = orc.emd_1d(x, tau=0.25, max_sifts=50, max_imf=1, end="natural")
    assert ns == st["per_mode_sifts"][0]
    assert np.max(np.abs(imf - o[0])) <= 1e-13 * np.max(np.abs(x))
