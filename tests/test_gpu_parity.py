"""GPU parity: the HIP path (through the C ABI) against the oracle.

FP64 mode uses the oracle's arithmetic, so extrema sets, IMF and sift counts
must match exactly and IMFs agree to rounding (tolerances written below).
Golden vectors from the reference's emd_ref.py pin the same cases end to end.
"""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from eemd4cuda_amd import _lib
    return _lib.engine()


def params(fp64=1, end="mirror", max_imf=8, max_sifts=10, tau=0.0, B=0, S=30, threads=0, mpt=1,
           seed=0, min_extrema=4):
    from eemd4cuda_amd import _lib
    p = _lib.default_params()
    p.fp64 = fp64
    p.end_mode = _lib.END_MODES[end]
    p.max_imf = max_imf
    p.max_sifts = max_sifts
    p.tau = tau
    p.block = B
    p.halo = S
    p.threads = threads
    p.members_per_task = mpt
    p.seed = seed
    p.min_extrema = min_extrema
    return p


def run1(eng, x, z, sigma, **kw):
    N = len(x)
    E = z.shape[0] if z is not None else kw.pop("E")
    prm = params(**kw)
    out = eng.run_batch(x, np.array([0, N]), E, np.array([sigma]), prm, noise=z,
                        member_imfs=True, trace=True)
    mi = prm.max_imf
    out["avg2"] = out["avg"].reshape(mi, N)[:out["n_modes"][0]]
    out["mimf"] = out["member_imfs"].reshape(E, mi, N)
    return out


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_fp64_three_tone_matches_oracle_and_reference(eng, golden):
    g = golden("golden_three_tone.npz")
    x, z, sigma = g["x"], g["z"], float(g["sigma"])
    print("case: three-tone N=4096 E=64 fp64 exact Thomas", flush=True)
    out = run1(eng, x, z, sigma, B=0)
    o = orc.eemd(x, z, sigma, tau=0.0, max_sifts=10, max_imf=8, end="mirror",
                 member_imfs=True, trace=True)
    assert np.array_equal(out["nmode"][0], o["nmode"])
    assert np.array_equal(out["nsift"][0], o["nsift"])
    assert np.array_equal(out["per_mode"][0], o["per_mode"])
    assert np.array_equal(out["trace"][0], o["trace"])        # extrema sets of every sift
    d = np.max(np.abs(out["mimf"] - o["member_imfs"]))
    print("member IMF max|d| vs oracle:", d)
    assert d <= 1e-13
    # against the reference's own output (golden)
    assert np.array_equal(out["nmode"][0], g["nmode"])
    assert np.array_equal(out["trace"][0], g["trace"])
    for m in range(len(g["avg"])):
        assert rel_l2(out["avg2"][m], g["avg"][m]) <= 1e-10
    assert np.max(np.abs(out["resid"] - g["resid"])) <= 1e-10


# B = 17 and 33: windows whose rows end inside a 16-row group, and (1024 threads) knot runs
# that cross many segment regions inside a group, for the solve's straight-line row groups
@pytest.mark.parametrize("B,S,threads,end", [(64, 30, 256, "mirror"), (100, 14, 1024, "mirror"),
                                             (1024, 30, 512, "mirror"), (256, 30, 64, "mirror"),
                                             (17, 5, 1024, "mirror"), (33, 30, 256, "natural")])
def test_fp64_dsbt_matches_oracle(eng, B, S, threads, end):
    rng = np.random.default_rng(7)
    N = 20000
    t = np.arange(N)
    x = np.sin(2 * np.pi * t / 1500) + 0.4 * np.sin(2 * np.pi * t / 97) + 0.1 * rng.standard_normal(N)
    z = rng.standard_normal((3, N))
    sigma = 0.2 * np.std(x)
    print("case: N=20000 E=3 fp64 DSBT B=%d S=%d threads=%d %s" % (B, S, threads, end), flush=True)
    out = run1(eng, x, z, sigma, B=B, S=S, threads=threads, max_imf=6, end=end)
    o = orc.eemd(x, z, sigma, tau=0.0, max_sifts=10, max_imf=6, end=end, B=B, S=S,
                 member_imfs=True, trace=True)
    assert np.array_equal(out["nmode"][0], o["nmode"])
    assert np.array_equal(out["trace"][0], o["trace"])
    assert np.max(np.abs(out["mimf"] - o["member_imfs"])) <= 1e-13
    assert np.max(np.abs(out["avg2"] - o["avg"])) <= 1e-13


@pytest.mark.parametrize("min_extrema,end", [(12, "mirror"), (40, "natural")])
def test_fp64_min_extrema_with_sd_stop_and_dsbt(eng, min_extrema, end):
    """A residual with fewer than min_extrema extrema ends the member
    (emd_ref.py:123-124), here with the SD stop (tau > 0) and DSBT windows: the
    IMF count, every sift's extrema and the IMFs equal the oracle's."""
    rng = np.random.default_rng(11)
    N = 6000
    t = np.arange(N)
    x = np.sin(2 * np.pi * t / 900) + 0.5 * np.sin(2 * np.pi * t / 61) + 0.05 * rng.standard_normal(N)
    z = rng.standard_normal((4, N))
    sigma = 0.2 * np.std(x)
    print("case: N=6000 E=4 fp64 tau=0.2 B=128 min_extrema=%d %s" % (min_extrema, end), flush=True)
    out = run1(eng, x, z, sigma, tau=0.2, max_sifts=30, max_imf=12, B=128, S=30, end=end,
               min_extrema=min_extrema)
    o = orc.eemd(x, z, sigma, tau=0.2, max_sifts=30, max_imf=12, min_extrema=min_extrema,
                 end=end, B=128, S=30, member_imfs=True, trace=True)
    assert np.array_equal(out["nmode"][0], o["nmode"])
    assert np.array_equal(out["nsift"][0], o["nsift"])
    assert np.array_equal(out["trace"][0], o["trace"])
    assert np.max(np.abs(out["mimf"] - o["member_imfs"])) <= 1e-13
    # the stop rule bites: fewer IMFs than with the default threshold
    assert int(out["n_modes"][0]) < 12


def test_fp64_natural_end_matches_reference(eng, golden):
    g = golden("golden_chirp_natural.npz")
    print("case: chirp N=2048 E=8 natural", flush=True)
    out = run1(eng, g["x"], g["z"], float(g["sigma"]), end="natural", max_imf=6, B=0)
    assert np.array_equal(out["nmode"][0], g["nmode"])
    assert np.array_equal(out["per_mode"][0], g["per_mode"])
    assert np.array_equal(out["trace"][0], g["trace"])
    assert np.max(np.abs(out["mimf"] - g["member_imfs"])) <= 1e-11
    assert np.max(np.abs(out["avg2"] - g["avg"])) <= 1e-11


@pytest.mark.parametrize("end,tau", [("natural", 0.25), ("mirror", 0.2)])
def test_fp64_sd_stopping_matches_reference(eng, golden, end, tau):
    g = golden("golden_chirp_emd.npz")
    k = "%s_tau" % end
    x = g["x"]
    print("case: chirp E=1 tau=%g %s" % (tau, end), flush=True)
    out = run1(eng, x, np.zeros((1, len(x))), 0.0, end=end, tau=tau, max_sifts=50,
               max_imf=64, B=0)
    nm = out["nmode"][0, 0]
    assert nm == len(g[k + "_per_mode"])
    assert np.array_equal(out["per_mode"][0, 0, :nm], g[k + "_per_mode"])
    assert np.array_equal(out["trace"][0, 0, :nm, :50], g[k + "_trace"])
    assert np.max(np.abs(out["mimf"][0, :nm] - g[k + "_imfs"])) <= 1e-10


def test_degenerate_inputs_match_reference(eng, golden):
    g = golden("golden_degenerate.npz")
    names = sorted(k[:-2] for k in g.files if k.endswith("_x"))
    for k in names:
        x = g[k + "_x"]
        end = "mirror" if k.endswith("mirror") else "natural"
        print("case: degenerate", k, flush=True)
        out = run1(eng, x, np.zeros((1, len(x))), 0.0, end=end, max_imf=6, B=0)
        nm = out["nmode"][0, 0]
        assert nm == g[k + "_imfs"].shape[0], k
        assert np.array_equal(out["per_mode"][0, 0, :nm], g[k + "_per_mode"]), k
        assert np.allclose(out["mimf"][0, :nm], g[k + "_imfs"], rtol=0, atol=1e-12), k


@pytest.mark.parametrize("end", ["natural", "mirror"])
def test_degenerate_inputs_in_a_one_wave_batch(eng, golden, end):
    """The same degenerate inputs (constant, monotone, plateaus, minima only,
    two knots, 8 samples, alternating) and the shortest legal signals (3 to 7
    samples) in one batched call with enough members (zero noise) to take one
    wave per member: every member equals the reference's golden EMD, and the
    short signals equal the oracle."""
    g = golden("golden_degenerate.npz")
    names = sorted(k[:-2] for k in g.files if k.endswith("_x") and k[:-2].endswith(end))
    rng = np.random.default_rng(5)
    shorts = [rng.standard_normal(n) for n in (3, 4, 5, 6, 7)]
    xs = [g[k + "_x"] for k in names] + shorts
    lens = np.array([len(v) for v in xs], dtype=np.int64)
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    E, mi = 180, 6
    print("case: degenerate + short signals in one batch, E=180, %s" % end, flush=True)
    out = eng.run_batch(np.concatenate(xs), offs, E, np.zeros(len(xs)),
                        params(end=end, max_imf=mi, B=0, mpt=0), member_imfs=True)
    assert eng.last_timing()["threads"] == 64
    mimf = out["member_imfs"]
    for s, v in enumerate(xs):
        n = len(v)
        got = mimf[E * mi * offs[s]: E * mi * offs[s + 1]].reshape(E, mi, n)
        if s < len(names):
            k = names[s]
            ref, pm = g[k + "_imfs"], g[k + "_per_mode"]
        else:
            o = orc.eemd(v, np.zeros((1, n)), 0.0, tau=0.0, max_sifts=10, max_imf=mi, end=end,
                         member_imfs=True)
            k = "short%d" % n
            ref, pm = o["member_imfs"][0, :o["nmode"][0]], o["per_mode"][0, :o["nmode"][0]]
        nm = ref.shape[0]
        assert np.all(out["nmode"][s] == nm), k
        assert np.all(out["per_mode"][s][:, :nm] == pm), k
        for e in (0, E // 2, E - 1):
            assert np.allclose(got[e, :nm], ref, rtol=0, atol=1e-12), (k, e)


def test_philox_noise_matches_oracle(eng):
    print("case: philox normals", flush=True)
    z = eng.noise(2024, 5, 3, 4, 1001)
    for e in range(4):
        ref = orc.normals(2024, 5, 3 + e, 1001)
        assert np.max(np.abs(z[e] - ref)) <= 1e-12


def test_in_kernel_noise_equals_materialised_noise(eng):
    x = np.sin(np.arange(6000) * 0.01) + 0.3 * np.sin(np.arange(6000) * 0.37)
    sigma = 0.2 * np.std(x)
    print("case: in-kernel philox vs materialised", flush=True)
    a = run1(eng, x, None, sigma, E=6, B=0, seed=99)
    z = eng.noise(99, 0, 0, 6, len(x))
    b = run1(eng, x, z, sigma, B=0)
    assert np.array_equal(a["nmode"], b["nmode"])
    assert np.array_equal(a["mimf"], b["mimf"])


def test_fp32_single_sift_accuracy(eng):
    """One sift in FP32 (no chaos yet): within FP32 rounding of the FP64 oracle."""
    from eemd4cuda_amd import signals
    x = signals.chirp2048()
    print("case: fp32 one sift", flush=True)
    for end in ("natural", "mirror"):
        out = run1(eng, x, np.zeros((1, len(x))), 0.0, fp64=0, end=end, max_imf=1, max_sifts=1,
                   B=1024, S=14)
        eu, _ = orc.envelope(x, True, end)
        el, _ = orc.envelope(x, False, end)
        h1 = x - 0.5 * (eu + el)
        err = np.max(np.abs(out["mimf"][0, 0] - h1)) / np.max(np.abs(x))
        print(end, "fp32 one-sift max rel err %.2e" % err)
        assert err <= 2e-6


@pytest.mark.parametrize("N,B", [(8192, 1024), (30000, 256), (3000, 1024)])
def test_fp32_window_solvers_one_sift(eng, N, B):
    """One FP32 sift with long block-Thomas windows (solved by a wave each
    when there are few, by a thread each otherwise) against the FP64 oracle
    with the same blocks: FP32 rounding only."""
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + np.sin(np.arange(N) * 0.01)
    print("case: fp32 one sift N=%d B=%d" % (N, B), flush=True)
    out = run1(eng, x, np.zeros((1, N)), 0.0, fp64=0, end="mirror", max_imf=1, max_sifts=1,
               B=B, S=14)
    eu, _ = orc.envelope(x, True, "mirror", B=B, S=14)
    el, _ = orc.envelope(x, False, "mirror", B=B, S=14)
    h1 = x - 0.5 * (eu + el)
    err = np.max(np.abs(out["mimf"][0, 0] - h1)) / np.max(np.abs(x))
    print("max rel err %.2e" % err)
    assert err <= 5e-6


def test_fp32_ensemble_error(eng, golden):
    """FP32 ensemble-mean IMFs vs the reference's FP64 output.  Extrema
    decisions between nearly equal neighbours flip under FP32 rounding and the
    sift amplifies each flip locally, so the error is reported per IMF; the
    bound here only guards against gross breakage."""
    g = golden("golden_three_tone.npz")
    x, z, sigma = g["x"], g["z"], float(g["sigma"])
    print("case: three-tone fp32", flush=True)
    out = run1(eng, x, z, sigma, fp64=0, B=1024, S=14)
    errs = [rel_l2(out["avg2"][m], g["avg"][m]) for m in range(min(len(g["avg"]), len(out["avg2"])))]
    print("fp32 per-IMF rel L2 vs reference:", ["%.2e" % e for e in errs])
    # measured 3.2e-3 4.9e-3 3.1e-4 9.2e-4 5.2e-4 2.2e-2 0.12 0.096: slow IMFs carry
    # the flips (a knot moved by one sample in a member changes its envelope
    # locally by O(1/spacing)); bounds just above the measured values
    bound = [1e-2, 1.5e-2, 1e-3, 3e-3, 2e-3, 6e-2, 0.3, 0.3]
    for m, b in enumerate(bound[:len(errs)]):
        assert errs[m] <= b, (m, errs[m])


def test_determinism(eng):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(9000).cumsum() * 0.05 + np.sin(np.arange(9000) * 0.2)
    print("case: determinism", flush=True)
    a = run1(eng, x, None, 0.1, E=12, B=256, S=30, seed=5, mpt=3)
    b = run1(eng, x, None, 0.1, E=12, B=256, S=30, seed=5, mpt=3)
    assert np.array_equal(a["avg"], b["avg"]) and np.array_equal(a["resid"], b["resid"])
    c = run1(eng, x, None, 0.1, E=12, B=256, S=30, seed=6, mpt=3)
    assert not np.array_equal(a["avg"], c["avg"])


def test_batched_uneven_equals_per_signal(eng):
    from eemd4cuda_amd import signals
    xs = signals.uneven_batch(5, lo_log2=9, hi_log2=13, seed=9)
    lens = [len(v) for v in xs]
    offs = np.concatenate(([0], np.cumsum(lens)))
    sig = np.array([0.2 * np.std(v) for v in xs])
    E = 5
    print("case: batched uneven", lens, flush=True)
    prm = params(B=128, S=30, max_imf=8)
    bat = eng.run_batch(np.concatenate(xs), offs, E, sig, prm, member_imfs=False)
    for s, v in enumerate(xs):
        one = eng.run_batch(v, np.array([0, len(v)]), E, sig[s:s + 1], params(B=128, S=30, max_imf=8))
        # Philox counters are keyed by the signal index: re-key the single run
        z = eng.noise(0, s, 0, E, len(v))
        one = eng.run_batch(v, np.array([0, len(v)]), E, sig[s:s + 1], params(B=128, S=30, max_imf=8),
                            noise=z)
        a = bat["avg"][8 * offs[s]: 8 * offs[s + 1]]
        assert np.array_equal(a, one["avg"]), s
        assert np.array_equal(bat["nmode"][s], one["nmode"][0]), s
        assert np.array_equal(bat["resid"][offs[s]:offs[s + 1]], one["resid"]), s


def test_workgroup_sizes_agree(eng):
    rng = np.random.default_rng(2)
    x = np.sin(np.arange(30000) * 0.003) + 0.2 * rng.standard_normal(30000)
    outs = []
    for nt in (256, 64, 512, 1024):
        print("case: threads", nt, flush=True)
        outs.append(run1(eng, x, None, 0.1, E=3, B=256, S=30, threads=nt))
    for o in outs[1:]:
        assert np.array_equal(o["mimf"], outs[0]["mimf"])


def test_members_per_task_only_reassociates(eng):
    x = np.sin(np.arange(5000) * 0.02) + 0.5 * np.sin(np.arange(5000) * 0.3)
    print("case: members per task", flush=True)
    a = run1(eng, x, None, 0.1, E=16, B=0, mpt=1)
    b = run1(eng, x, None, 0.1, E=16, B=0, mpt=5)
    assert np.array_equal(a["mimf"], b["mimf"])
    assert np.max(np.abs(a["avg"] - b["avg"])) <= 1e-13


def _sd_case(N):
    rng = np.random.default_rng(N)
    t = np.arange(N)
    x = np.sin(2 * np.pi * t / 50.0) + 0.5 * np.sin(2 * np.pi * t / 7.3) + 0.1 * rng.standard_normal(N)
    # h after sifts 1 and 2 of the first IMF (tau = 0: exactly that many sifts)
    h1 = orc.emd_1d(x, tau=0.0, max_sifts=1, max_imf=1, end="mirror")[0][0]
    h2 = orc.emd_1d(x, tau=0.0, max_sifts=2, max_imf=1, end="mirror")[0][0]
    sd = float(np.sum((h1 - h2) ** 2) / np.sum(h1 ** 2))   # emd_ref.py:141-143
    return x, sd


@pytest.mark.parametrize("N", [1000, 8192, 20000, 40961])
def test_sd_stop_decision_at_one_ulp(eng, N):
    """tau set to the reference's SD of the second sift, and one ulp either
    side: the stop decision hinges on the last bit of the SD ratio, so the
    sums must be added in np.sum's pairwise order (blocks of 8192, pairwise
    inside, leaves of 128 with eight running sums)."""
    x, sd = _sd_case(N)
    for tau, stops in ((sd, False), (np.nextafter(sd, np.inf), True), (np.nextafter(sd, -np.inf), False)):
        print("case: SD stop N=%d tau=%r" % (N, tau), flush=True)
        out = run1(eng, x, np.zeros((1, N)), 0.0, end="mirror", tau=tau, max_sifts=50, max_imf=1,
                   B=0)
        _, _, st = orc.emd_1d(x, tau=tau, max_sifts=50, max_imf=1, end="mirror")
        ns = int(out["per_mode"][0, 0, 0])
        assert ns == st["per_mode_sifts"][0]
        assert (ns == 2) == stops, (tau, ns)
