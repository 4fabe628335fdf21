"""The CPU oracle, pinned independently of itself.

Pins (each chosen so a plausible slip -- a dropped term, a wrong sign, an
off-by-one interval, a transposed coupling -- fails one of them):
* golden vectors written by the reference's own emd_ref.py
  (tests/golden/make_golden.py): per-member IMF/sift counts, per-sift
  extrema digests, IMFs, ensemble means, residuals;
* extrema vs a brute-force loop and the reference's U2 edge cases
  (audit_eemd/test_block_thomas.py:211-238, re-authored);
* the spline system + Thomas vs dense numpy.linalg.solve (emd_ref.py:264-278);
* natural / mirror envelopes vs scipy.interpolate.CubicSpline(bc_type='natural');
* DSBT vs exact Thomas within the decay bound C*rho^S, rho = 2 - sqrt(3);
* Philox4x32-10 vs the published Random123 known-answer vectors;
* Box-Muller vs its closed form on the same uniforms; moments;
* pairwise summation vs numpy.sum.
"""
import numpy as np
import pytest

from oracle import emd_np, orc


# ---------------------------------------------------------------- extrema
def _brute(x, want_max):
    out = []
    for i in range(1, len(x) - 1):
        if want_max and x[i] > x[i - 1] and x[i] > x[i + 1]:
            out.append(i)
        if not want_max and x[i] < x[i - 1] and x[i] < x[i + 1]:
            out.append(i)
    return np.array(out, dtype=np.int64)


def _scan_cases():
    rng = np.random.default_rng(0)
    c = {"random": rng.standard_normal(5000)}
    x = rng.standard_normal(4000); x[100:110] = x[100]; x[200:260] = x[200]
    c["flat_runs"] = x
    x = rng.standard_normal(4000); x[500:505] = 99.0
    c["plateau"] = x
    x = np.linspace(0, 1, 1000); x[500] = 5.0
    c["single_max"] = x
    x = np.linspace(0, 1, 1000); x[500] = -5.0
    c["single_min"] = x
    c["constant"] = np.ones(500)
    c["monotone"] = np.linspace(-3, 3, 2000)
    c["alternating"] = (np.arange(1000) % 2) * 2.0 - 1.0
    c["tiny"] = np.array([0.0, 1.0, 0.0])
    return c


@pytest.mark.parametrize("name", list(_scan_cases()))
def test_extrema_vs_brute_force(name):
    x = _scan_cases()[name]
    for want in (True, False):
        assert np.array_equal(orc.extrema(x, want), _brute(x, want))


# ---------------------------------------------------------------- Thomas
def _system(T, Y):
    H = np.diff(T)
    n = len(T) - 2
    d = 2.0 * (H[:n] + H[1:n + 1])
    l = np.zeros(n); u = np.zeros(n)
    l[1:] = H[1:n]; u[:n - 1] = H[1:n]
    r = 6.0 * ((Y[2:] - Y[1:-1]) / H[1:] - (Y[1:-1] - Y[:-2]) / H[:-1])
    return d, l, u, r


def _dense(d, l, u):
    n = len(d)
    A = np.diag(d)
    A[np.arange(n - 1), np.arange(1, n)] = u[:n - 1]
    A[np.arange(1, n), np.arange(n - 1)] = l[1:]
    return A


@pytest.mark.parametrize("kind", ["uniform", "cv04", "alt110", "cv15"])
def test_thomas_vs_dense(kind):
    rng = np.random.default_rng(1)
    p = 600
    if kind == "uniform":
        T = np.arange(p) * 10.0
    elif kind == "cv04":
        T = np.concatenate(([0.0], np.cumsum(np.clip(10 * (1 + 0.4 * rng.standard_normal(p - 1)), 1, None))))
    elif kind == "alt110":
        T = np.concatenate(([0.0], np.cumsum(np.resize([1.0, 10.0], p - 1))))
    else:
        T = np.concatenate(([0.0], np.cumsum(np.clip(10 * np.exp(0.75 * rng.standard_normal(p - 1)), 1, None))))
    T = np.round(T)
    T = np.maximum.accumulate(T + np.arange(p) * 1e-9).round()
    for i in range(1, p):
        if T[i] <= T[i - 1]:
            T[i] = T[i - 1] + 1
    Y = rng.standard_normal(p)
    d, l, u, r = _system(T, Y)
    x_dense = np.linalg.solve(_dense(d, l, u), r)
    x = orc.dsbt(d, l, u, r)              # exact Thomas
    assert np.max(np.abs(x - x_dense)) <= 1e-13 * np.max(np.abs(x_dense))
    # decay-seamed block Thomas: error <= C rho^S (rho = 2 - sqrt(3))
    rho = 2 - np.sqrt(3)
    for B, S in ((64, 8), (64, 14), (100, 30)):
        xs = orc.dsbt(d, l, u, r, B=B, S=S)
        err = np.max(np.abs(xs - x_dense)) / np.max(np.abs(x_dense))
        assert err <= 8 * rho ** S + 1e-15, (B, S, err)
    # one window covering everything is exact Thomas, bit for bit
    assert np.array_equal(orc.dsbt(d, l, u, r, B=len(d), S=30), x)


def test_thomas_small_n():
    d = np.array([4.0]); l = np.zeros(1); u = np.zeros(1); r = np.array([2.0])
    assert orc.dsbt(d, l, u, r)[0] == 0.5
    assert len(orc.dsbt(np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0))) == 0


# ---------------------------------------------------------------- envelopes
@pytest.mark.parametrize("end", ["natural", "mirror"])
@pytest.mark.parametrize("want", [True, False])
def test_envelope_vs_scipy(end, want):
    from scipy.interpolate import CubicSpline
    x = np.load(__import__("tests.conftest", fromlist=["GOLDEN"]).GOLDEN / "golden_chirp_emd.npz")["x"]
    env, p = orc.envelope(x, want, end)
    k = _brute(x, want)
    T = k.astype(float); Y = x[k]
    if end == "mirror" and p >= 3:
        T = np.concatenate(([T[0] - (T[1] - T[0])], T, [T[-1] + (T[-1] - T[-2])]))
        Y = np.concatenate(([Y[1]], Y, [Y[-2]]))
    cs = CubicSpline(T, Y, bc_type="natural", extrapolate=True)
    ref = cs(np.arange(len(x), dtype=float))
    assert np.max(np.abs(env - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_envelope_interpolates_knots_exactly():
    x = np.load(__import__("tests.conftest", fromlist=["GOLDEN"]).GOLDEN / "golden_chirp_emd.npz")["x"]
    for end in ("natural", "mirror"):
        env, _ = orc.envelope(x, True, end)
        k = _brute(x, True)
        assert np.array_equal(env[k[1:]], x[k[1:]])  # left-closed intervals give y exactly


def test_envelope_degenerate_counts():
    x = np.zeros(20); x[7] = 1.0
    env, p = orc.envelope(x, True)
    assert p == 1 and np.all(env == 1.0)
    env, p = orc.envelope(np.linspace(0, 1, 20), True)
    assert p == 0 and np.all(env == 0.0)


# ---------------------------------------------------------------- golden vectors
def test_eemd_three_tone_vs_reference(golden):
    g = golden("golden_three_tone.npz")
    r = orc.eemd(g["x"], g["z"], float(g["sigma"]), tau=0.0, max_sifts=10, max_imf=8,
                 end="mirror", member_imfs=True, trace=True)
    assert np.array_equal(r["nmode"], g["nmode"])
    assert np.array_equal(r["nsift"], g["nsift"])
    assert np.array_equal(r["per_mode"], g["per_mode"])
    assert np.array_equal(r["trace"], g["trace"])           # every sift's extrema sets
    for m in range(len(g["avg"])):
        rel = np.linalg.norm(r["avg"][m] - g["avg"][m]) / np.linalg.norm(g["avg"][m])
        assert rel <= 1e-12, (m, rel)
    assert np.max(np.abs(r["resid"] - g["resid"])) <= 1e-12
    assert np.max(np.abs(r["member_imfs"][:4] - g["member_imfs_0_3"])) <= 1e-12


def test_eemd_natural_vs_reference(golden):
    g = golden("golden_chirp_natural.npz")
    r = orc.eemd(g["x"], g["z"], float(g["sigma"]), tau=0.0, max_sifts=10, max_imf=6,
                 end="natural", member_imfs=True, trace=True)
    assert np.array_equal(r["nmode"], g["nmode"])
    assert np.array_equal(r["per_mode"], g["per_mode"])
    assert np.array_equal(r["trace"], g["trace"])
    assert np.max(np.abs(r["member_imfs"] - g["member_imfs"])) <= 1e-12
    assert np.max(np.abs(r["avg"] - g["avg"])) <= 1e-12


@pytest.mark.parametrize("end,tau", [("natural", 0.25), ("mirror", 0.2)])
def test_emd_sd_stop_vs_reference(golden, end, tau):
    g = golden("golden_chirp_emd.npz")
    k = "%s_tau" % end
    imfs, resid, st = orc.emd_1d(g["x"], tau=tau, max_sifts=50, max_imf=-1, end=end, trace=True)
    assert st["per_mode_sifts"] == list(g[k + "_per_mode"])
    assert np.array_equal(st["trace"], g[k + "_trace"])
    assert np.max(np.abs(imfs - g[k + "_imfs"])) <= 1e-11
    assert np.max(np.abs(resid - g[k + "_resid"])) <= 1e-11


def test_envelopes_vs_reference(golden):
    g = golden("golden_chirp_emd.npz")
    for end in ("natural", "mirror"):
        for want in ("max", "min"):
            env, _ = orc.envelope(g["x"], want == "max", end)
            ref = g["env_%s_%s" % (want, end)]
            assert np.max(np.abs(env - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_degenerate_vs_reference(golden):
    g = golden("golden_degenerate.npz")
    names = sorted(k[:-2] for k in g.files if k.endswith("_x"))
    assert names
    for k in names:
        x = g[k + "_x"]
        end = "mirror" if k.endswith("mirror") else "natural"
        imfs, resid, st = orc.emd_1d(x, tau=0.0, max_sifts=10, max_imf=6, end=end)
        assert st["per_mode_sifts"] == list(g[k + "_per_mode"]), k
        assert imfs.shape == g[k + "_imfs"].shape, k
        assert np.allclose(imfs, g[k + "_imfs"], rtol=0, atol=1e-12), k
        assert np.allclose(resid, g[k + "_resid"], rtol=0, atol=1e-12), k


def test_numpy_port_is_the_reference(golden):
    """The timed CPU baseline reproduces emd_ref bit for bit."""
    g = golden("golden_three_tone.npz")
    for e in range(2):
        imfs, _, st = emd_np.emd_1d(g["x"] + g["z"][e] * g["sigma"], tau=0.0, max_sifts=10,
                                    max_imf=8, end="mirror")
        assert np.array_equal(np.array(imfs), g["member_imfs_0_3"][e][:g["nmode"][e]])
    g = golden("golden_chirp_emd.npz")
    imfs, resid, st = emd_np.emd_1d(g["x"], tau=0.25, max_sifts=50, end="natural")
    assert np.array_equal(np.array(imfs), g["natural_tau_imfs"])


def test_partition_of_unity():
    rng = np.random.default_rng(3)
    x = np.cumsum(rng.standard_normal(3000)) * 0.1 + np.sin(np.arange(3000) * 0.05)
    imfs, resid, st = orc.emd_1d(x, tau=0.0, max_sifts=10, max_imf=-1, end="mirror")
    assert np.max(np.abs(x - (imfs.sum(0) + resid))) <= 1e-12
    assert st["total_sifts"] == 10 * st["n_modes"] or st["total_sifts"] < 10 * st["n_modes"]


def test_dc_invariance():
    """Adding a constant shifts only the residual (envelopes are translation
    equivariant): the IMFs are unchanged up to rounding."""
    x = np.load(__import__("tests.conftest", fromlist=["GOLDEN"]).GOLDEN / "golden_chirp_emd.npz")["x"]
    a, ra, _ = orc.emd_1d(x, tau=0.0, max_sifts=10, max_imf=3, end="mirror")
    b, rb, _ = orc.emd_1d(x + 0.25, tau=0.0, max_sifts=10, max_imf=3, end="mirror")
    assert a.shape == b.shape
    assert np.max(np.abs(a - b)) <= 1e-9
    assert np.max(np.abs((rb - ra) - 0.25)) <= 1e-9


# ---------------------------------------------------------------- noise
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert list(orc.philox4x32_10([0, 0, 0, 0], [0, 0])) == \
        [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(orc.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2)) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(orc.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                  [0xa4093822, 0x299f31d0])) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_box_muller_closed_form():
    seed, sig, mem, N = 12345, 3, 7, 64
    z = orc.normals(seed, sig, mem, N)
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for j in range(N // 2):
        w = [int(v) for v in orc.philox4x32_10([j, mem, sig, 0x45454D44], key)]
        a = ((w[0] << 32) | w[1]) >> 11
        b = ((w[2] << 32) | w[3]) >> 11
        u1 = (a + 1) * 2.0 ** -53
        u2 = b * 2.0 ** -53
        rad = np.sqrt(-2 * np.log(u1))
        assert abs(z[2 * j] - rad * np.cos(2 * np.pi * u2)) <= 1e-13
        assert abs(z[2 * j + 1] - rad * np.sin(2 * np.pi * u2)) <= 1e-13


def test_normal_moments_and_independence():
    z = orc.normals(0, 0, 0, 400000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert abs(np.mean(z ** 3)) < 0.03 and abs(np.mean(z ** 4) - 3) < 0.05
    z2 = orc.normals(0, 0, 1, 400000)
    assert abs(np.corrcoef(z, z2)[0, 1]) < 0.01


def test_pairwise_sum_is_numpy_sum():
    rng = np.random.default_rng(4)
    for n in (1, 7, 8, 9, 127, 128, 129, 1000, 4097, 100003):
        a = rng.standard_normal(n) * 10.0 ** rng.uniform(-8, 8, n)
        assert orc.pairwise_sum(a) == np.sum(a)


@pytest.mark.parametrize("N", [1000, 20000])
def test_oracle_sd_stop_decision_at_one_ulp_equals_numpy(N):
    """The oracle's SD stop, at tau one ulp either side of the SD of the
    second sift, decides as the numpy restatement of emd_ref (np.sum)."""
    from oracle import emd_np
    rng = np.random.default_rng(N)
    t = np.arange(N)
    x = np.sin(2 * np.pi * t / 50.0) + 0.5 * np.sin(2 * np.pi * t / 7.3) + 0.1 * rng.standard_normal(N)
    h1 = orc.emd_1d(x, tau=0.0, max_sifts=1, max_imf=1, end="mirror")[0][0]
    h2 = orc.emd_1d(x, tau=0.0, max_sifts=2, max_imf=1, end="mirror")[0][0]
    sd = float(np.sum((h1 - h2) ** 2) / np.sum(h1 ** 2))
    for tau in (sd, np.nextafter(sd, np.inf), np.nextafter(sd, -np.inf)):
        _, _, st = orc.emd_1d(x, tau=tau, max_sifts=50, max_imf=1, end="mirror")
        _, _, sn = emd_np.emd_1d(x, tau=tau, max_sifts=50, max_imf=1, end="mirror")
        assert st["per_mode_sifts"] == sn["per_mode_sifts"]
    assert st["per_mode_sifts"][0] > 2
