"""GPU at a large size: one member of 2^23 + 12345 samples (wave segments of
over a million samples, 64-bit offsets, knot regions of millions of slots, a
ragged last chunk).  FP64 against the oracle bit for bit (extrema of every
sift, IMF values to rounding) on 2 IMFs x 3 sifts with block-Thomas windows;
FP32 one sift within FP32 rounding of the oracle's envelopes."""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu

N = (1 << 23) + 12345


@pytest.fixture(scope="module")
def eng():
    from eemd4cuda_amd import _lib
    return _lib.engine()


@pytest.fixture(scope="module")
def x():
    rng = np.random.default_rng(11)
    t = np.arange(N)
    return (np.sin(2 * np.pi * t / 5000.0) + 0.3 * np.sin(2 * np.pi * t / 37.0)
            + 0.05 * rng.standard_normal(N))


def _params(fp64, max_imf, max_sifts, S):
    from eemd4cuda_amd import _lib
    p = _lib.default_params()
    p.fp64 = fp64
    p.end_mode = _lib.END_MODES["mirror"]
    p.max_imf = max_imf
    p.max_sifts = max_sifts
    p.block = 1024
    p.halo = S
    p.members_per_task = 1
    return p


def test_fp64_large_member_matches_oracle(eng, x):
    z = np.zeros((1, N))
    print("case: fp64 N=%d 2 IMFs x 3 sifts B=1024" % N, flush=True)
    out = eng.run_batch(x, np.array([0, N]), 1, np.array([0.0]), _params(1, 2, 3, 30), noise=z,
                        member_imfs=True, trace=True)
    o = orc.eemd(x, z, 0.0, tau=0.0, max_sifts=3, max_imf=2, end="mirror", B=1024, S=30,
                 member_imfs=True, trace=True)
    assert np.array_equal(out["nmode"][0], o["nmode"])
    assert np.array_equal(out["trace"][0], o["trace"])
    d = np.max(np.abs(out["member_imfs"].reshape(1, 2, N) - o["member_imfs"]))
    print("member IMF max|d| vs oracle:", d)
    assert d <= 1e-12
    # partition of unity of the member: IMFs + residual = x
    assert np.max(np.abs(out["member_imfs"].reshape(2, N).sum(axis=0) + out["resid"] - x)) <= 1e-12


def test_fp32_large_member_one_sift(eng, x):
    """With 8.4M noisy samples some neighbours differ by less than FP32
    resolution, and rounding the input to float decides differently whether
    they are extrema (a local envelope change, not an arithmetic error).  The
    oracle therefore gets the same float-rounded input, so both take identical
    extrema and the comparison measures the FP32 envelope arithmetic alone."""
    x32 = x.astype(np.float32).astype(np.float64)
    print("case: fp32 N=%d one sift" % N, flush=True)
    out = eng.run_batch(x32, np.array([0, N]), 1, np.array([0.0]), _params(0, 1, 1, 14),
                        noise=np.zeros((1, N)), member_imfs=True)
    eu, _ = orc.envelope(x32, True, "mirror", B=1024, S=14)
    el, _ = orc.envelope(x32, False, "mirror", B=1024, S=14)
    h1 = x32 - 0.5 * (eu + el)
    err = np.max(np.abs(out["member_imfs"][:N] - h1)) / np.max(np.abs(x32))
    print("fp32 one-sift max rel err %.2e" % err)
    assert err <= 5e-6
    # the same sift from the unrounded input differs only near FP32 near-ties:
    # a few short clusters of samples
    eu, _ = orc.envelope(x, True, "mirror", B=1024, S=14)
    el, _ = orc.envelope(x, False, "mirror", B=1024, S=14)
    d = np.abs(out["member_imfs"][:N] - (x - 0.5 * (eu + el)))
    print("samples off by > 1e-4 against the unrounded input:", int(np.sum(d > 1e-4)))
    assert np.sum(d > 1e-4) <= 1000
