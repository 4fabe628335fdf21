"""GPU parity at the benchmark configurations' full signal lengths.

* audio_mix: one 48 kHz record of 2^20 samples, 10 IMFs x 10 sifts,
  decay-seamed block-Thomas B=1024 / S=30, in-kernel Philox noise.  Two
  members in FP64 equal the oracle (extrema traces, IMF and sift counts bit
  for bit; IMFs to rounding).  All 256 members: partition of unity, fixed
  sift counts, and the FP32 error per ensemble-mean IMF against FP64, bounded
  just above its measured value.
* uneven: the 512-signal batch (lengths 2^12 .. 2^18, seed 2), E=100, 8 IMFs,
  in one batched call.  Member traces of 8 signals x 2 members and the
  ensemble means of 3 signals equal the oracle (FP64); every signal satisfies
  the partition of unity; FP32 error per IMF against FP64 is bounded.
* chirp/AM: the first 4 of the 4096 x 65536 set (seed 4) with learned ends
  (packaged model) in FP64 equal the oracle with the same model.
* the reference's first-IMF EEMD gpu_eemd (emd_gpu.py:337-358) equals the
  oracle's first-IMF member mean.
* 300 short signals of ragged lengths with natural ends in one call (one wave
  per member): picked members equal the oracle bit for bit.
* the automatic members-per-task count is one for the whole batch and only
  regroups the ensemble sum.
* the SD stop (tau > 0) in a one-wave-per-member batch with DSBT windows:
  picked members equal the oracle bit for bit.
"""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from eemd4cuda_amd import _lib
    return _lib.engine()


def params(fp64=1, end="mirror", max_imf=10, max_sifts=10, B=1024, S=30, seed=1, mpt=1):
    from eemd4cuda_amd import _lib
    p = _lib.default_params()
    p.fp64 = fp64
    p.end_mode = _lib.END_MODES[end]
    p.max_imf = max_imf
    p.max_sifts = max_sifts
    p.block = B
    p.halo = S
    p.seed = seed
    p.members_per_task = mpt
    return p


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def means(out, lens, mi):
    avg = out["avg"]
    offs = np.concatenate(([0], np.cumsum(lens)))
    return [avg[mi * offs[s]: mi * offs[s + 1]].reshape(mi, lens[s])[:out["n_modes"][s]]
            for s in range(len(lens))]


@pytest.fixture(scope="module")
def audio():
    from eemd4cuda_amd import signals
    return signals.audio_mix(1 << 20, seed=1)


def test_audio_mix_fp64_two_members_match_oracle(eng, audio):
    x = audio
    N = len(x)
    sigma = 0.2 * np.std(x)
    E = 2
    print("case: audio_mix N=2^20 E=2 fp64 B=1024 S=30 philox seed 1", flush=True)
    out = eng.run_batch(x, np.array([0, N]), E, np.array([sigma]), params(), member_imfs=True,
                        trace=True)
    z = np.stack([orc.normals(1, 0, e, N) for e in range(E)])
    o = orc.eemd(x, z, sigma, tau=0.0, max_sifts=10, max_imf=10, end="mirror", B=1024, S=30,
                 member_imfs=True, trace=True)
    assert np.array_equal(out["nmode"][0], o["nmode"])
    assert np.array_equal(out["nsift"][0], o["nsift"])
    assert np.array_equal(out["trace"][0], o["trace"])
    mimf = out["member_imfs"].reshape(E, 10, N)
    for e in range(E):
        for m in range(int(o["nmode"][e])):
            assert rel_l2(mimf[e, m], o["member_imfs"][e, m]) <= 1e-10, (e, m)
    d = float(np.max(np.abs(mimf - o["member_imfs"])))
    print("member IMF max|d| vs oracle %.2e" % d)
    assert d <= 1e-12 * np.max(np.abs(x))
    avg = out["avg"].reshape(10, N)[:out["n_modes"][0]]
    for m in range(o["n_modes"]):
        assert rel_l2(avg[m], o["avg"][m]) <= 1e-10


def test_audio_mix_full_ensemble_invariants_and_fp32_error(eng, audio):
    x = audio
    N = len(x)
    sigma = 0.2 * np.std(x)
    E = 256
    print("case: audio_mix N=2^20 E=256 fp64 and fp32", flush=True)
    o64 = eng.run_batch(x, np.array([0, N]), E, np.array([sigma]), params(mpt=0))
    o32 = eng.run_batch(x, np.array([0, N]), E, np.array([sigma]), params(fp64=0, S=14, mpt=0))
    for o in (o64, o32):
        nm = o["n_modes"][0]
        avg = o["avg"].reshape(10, N)[:nm]
        # x = sum of the mean IMFs + residual (emd_ref.py:217)
        assert np.max(np.abs(avg.sum(axis=0) + o["resid"] - x)) <= 1e-12 * np.max(np.abs(x))
        # tau = 0: every IMF takes exactly max_sifts sifts
        assert np.array_equal(o["nsift"][0], 10 * o["nmode"][0])
        assert np.all(o["nmode"][0] == 10)
    a64 = o64["avg"].reshape(10, N)
    a32 = o32["avg"].reshape(10, N)
    err = [rel_l2(a32[m], a64[m]) for m in range(10)]
    err_x = [float(np.linalg.norm(a32[m] - a64[m]) / np.linalg.norm(x)) for m in range(10)]
    print("fp32 rel L2 per IMF vs fp64:", ["%.2e" % e for e in err])
    print("fp32 L2 per IMF vs fp64 / |x|:", ["%.2e" % e for e in err_x])
    # measured 7.5e-5 1.5e-4 2.9e-4 5.2e-4 1.9e-3 3.0e-3 2.6e-3 3.5e-2 (IMFs 1-8) and, for the
    # two slowest levels where members disagree on their existence, 2.7e-2 and 1.2 of |x|
    bound = [2.5e-4, 5e-4, 1e-3, 2e-3, 6e-3, 1e-2, 1e-2, 0.1]
    for m, b in enumerate(bound):
        assert err[m] <= b, (m, err[m])
    assert err_x[8] <= 0.1 and err_x[9] <= 4.0


@pytest.fixture(scope="module")
def uneven():
    from eemd4cuda_amd import signals
    xs = signals.uneven_batch(512, seed=2)
    return xs


def test_uneven_batch_fp64_matches_oracle(eng, uneven):
    xs = uneven
    lens = np.array([len(v) for v in xs])
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    sig = np.array([0.2 * np.std(v) for v in xs])
    E, mi = 100, 8
    print("case: uneven 512 signals E=100 fp64 one batched call, trace", flush=True)
    out = eng.run_batch(np.concatenate(xs), offs, E, sig, params(max_imf=mi, mpt=0, seed=2),
                        trace=True)
    xcat = np.concatenate(xs)
    ms = means(out, lens, mi)
    for s in range(len(xs)):
        r = out["resid"][offs[s]:offs[s + 1]]
        assert np.max(np.abs(ms[s].sum(axis=0) + r - xs[s])) <= 1e-12 * np.max(np.abs(xs[s])), s
    # per-mode sift counts add up, and no IMF takes more than max_sifts sifts
    assert np.array_equal(out["per_mode"].sum(axis=2), out["nsift"])
    assert out["per_mode"].max() <= 10
    order = np.argsort(lens)
    pick = order[np.linspace(0, len(xs) - 1, 8).astype(int)]   # shortest .. longest
    for s in pick:
        n = int(lens[s])
        z = np.stack([orc.normals(2, int(s), e, n) for e in range(2)])
        o = orc.eemd(xs[s], z, sig[s], tau=0.0, max_sifts=10, max_imf=mi, end="mirror", B=1024,
                     S=30, trace=True)
        assert np.array_equal(out["nmode"][s, :2], o["nmode"]), s
        assert np.array_equal(out["nsift"][s, :2], o["nsift"]), s
        assert np.array_equal(out["trace"][s, :2], o["trace"]), s
    for s in pick[[0, 4, 6]]:
        n = int(lens[s])
        print("oracle ensemble mean of signal %d (n=%d, E=100)" % (s, n), flush=True)
        z = orc.normals_matrix(2, int(s), E, n)
        o = orc.eemd(xs[s], z, sig[s], tau=0.0, max_sifts=10, max_imf=mi, end="mirror", B=1024,
                     S=30)
        assert np.array_equal(out["nmode"][s], o["nmode"]), s
        assert len(ms[s]) == o["n_modes"]
        for m in range(o["n_modes"]):
            assert rel_l2(ms[s][m], o["avg"][m]) <= 1e-10, (s, m)
    del xcat


def test_uneven_batch_fp32_error_per_imf(eng, uneven):
    xs = uneven
    lens = np.array([len(v) for v in xs])
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    sig = np.array([0.2 * np.std(v) for v in xs])
    E, mi = 100, 8
    print("case: uneven 512 signals E=100 fp32 vs fp64 (automatic tasks)", flush=True)
    o64 = eng.run_batch(np.concatenate(xs), offs, E, sig, params(max_imf=mi, mpt=0, seed=2))
    o32 = eng.run_batch(np.concatenate(xs), offs, E, sig,
                        params(fp64=0, S=14, max_imf=mi, mpt=0, seed=2))
    m64, m32 = means(o64, lens, mi), means(o32, lens, mi)
    worst = np.zeros(mi)
    for s in range(len(xs)):
        r = o32["resid"][offs[s]:offs[s + 1]]
        assert np.max(np.abs(m32[s].sum(axis=0) + r - xs[s])) <= 1e-6 * np.max(np.abs(xs[s]))
        xn = np.linalg.norm(xs[s])
        for m in range(min(len(m32[s]), len(m64[s]))):
            worst[m] = max(worst[m], float(np.linalg.norm(m32[s][m] - m64[s][m]) / xn))
    print("fp32 worst-signal L2 per IMF vs fp64 / |x|:", ["%.2e" % e for e in worst])
    # measured 2.4e-4 6.0e-4 9.5e-4 2.8e-3 of |x| for IMFs 1-4 (later levels: members
    # disagree on which levels exist, 2.5e-2 .. 0.55)
    assert np.all(worst[:4] <= np.array([1e-3, 2e-3, 3e-3, 1e-2])), worst


def test_chirp_learned_fp64_matches_oracle(eng):
    from eemd4cuda_amd import learned, signals
    xs = signals.chirp_am_batch(4, 65536, seed=4)
    model = learned.load(learned.default_model_path())
    learned.install(model, eng)
    N, E = xs.shape[1], 4
    offs = np.arange(5, dtype=np.int64) * N
    sig = 0.2 * np.std(xs, axis=1)
    print("case: chirp_am 4 x 65536 E=4 fp64 learned ends", flush=True)
    p = params(end="learned", seed=4)
    out = eng.run_batch(xs.ravel(), offs, E, sig, p, trace=True, member_imfs=True)
    mimf = out["member_imfs"].reshape(4, E, 10, N)
    orc.set_model(model)
    try:
        for s in range(4):
            z = orc.normals_matrix(4, s, E, N)
            o = orc.eemd(xs[s], z, sig[s], tau=0.0, max_sifts=10, max_imf=10, end="learned",
                         B=1024, S=30, member_imfs=True, trace=True)
            assert np.array_equal(out["nmode"][s], o["nmode"]), s
            assert np.array_equal(out["trace"][s], o["trace"]), s
            for e in range(E):
                for m in range(int(o["nmode"][e])):
                    assert rel_l2(mimf[s, e, m], o["member_imfs"][e, m]) <= 1e-10, (s, e, m)
    finally:
        orc.set_model(None)


def test_gpu_eemd_first_imf_matches_oracle():
    """The legacy first-IMF EEMD (emd_gpu.py:337-358): numpy noise from
    default_rng(seed), one IMF per member, mean over all E members."""
    from eemd4cuda_amd import signals
    from eemd4cuda_amd.emd_gpu import gpu_eemd
    x = signals.chirp2048()
    E = 8
    print("case: gpu_eemd chirp2048 E=8 tau=0.25", flush=True)
    avg, nsift, raw = gpu_eemd(x, E=E, eps=0.2, tau=0.25, max_sifts=50, seed=0, B=250, S=30)
    z = np.random.default_rng(0).standard_normal((E, len(x)))
    o = orc.eemd(x, z, 0.2 * np.std(x), tau=0.25, max_sifts=50, max_imf=1, end="natural",
                 B=250, S=30, member_imfs=True)
    assert np.array_equal(nsift, o["nsift"])
    assert np.max(np.abs(raw - o["member_imfs"][:, 0])) <= 1e-12
    assert np.max(np.abs(avg - o["member_imfs"][:, 0].mean(axis=0))) <= 1e-12


def test_short_ragged_batch_natural_ends_matches_oracle(eng):
    """Hundreds of short signals of ragged lengths (40 .. 3000 samples, most
    shorter than a wave's 512-sample chunk) in one call with natural ends: the
    launch takes one wave per member, and a picked member's extrema of every
    sift, IMF and sift counts and IMFs equal the oracle's (FP64, Philox noise,
    exact Thomas)."""
    rng = np.random.default_rng(21)
    lens = rng.integers(40, 3000, size=300)
    xs = []
    for n in lens:
        t = np.arange(n)
        f1, f2 = rng.uniform(0.005, 0.2, size=2)
        xs.append(np.sin(2 * np.pi * f1 * t) + 0.4 * np.sin(2 * np.pi * f2 * t + 1.0)
                  + 0.05 * rng.standard_normal(n))
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    sig = np.array([0.2 * np.std(v) for v in xs])
    E, mi = 8, 6
    print("case: 300 short ragged signals E=8 fp64 natural, one call", flush=True)
    out = eng.run_batch(np.concatenate(xs), offs, E, sig,
                        params(max_imf=mi, end="natural", B=0, mpt=0, seed=7),
                        member_imfs=True, trace=True)
    assert eng.last_timing()["threads"] == 64
    ms = means(out, lens, mi)
    for s in range(len(xs)):
        r = out["resid"][offs[s]:offs[s + 1]]
        assert np.max(np.abs(ms[s].sum(axis=0) + r - xs[s])) <= 1e-12 * np.max(np.abs(xs[s])), s
    order = np.argsort(lens)
    pick = order[np.linspace(0, len(xs) - 1, 10).astype(int)]
    mimf = out["member_imfs"]
    for s in pick:
        n = int(lens[s])
        z = np.stack([orc.normals(7, int(s), e, n) for e in range(E)])
        o = orc.eemd(xs[s], z, sig[s], tau=0.0, max_sifts=10, max_imf=mi, end="natural", B=0,
                     member_imfs=True, trace=True)
        assert np.array_equal(out["nmode"][s], o["nmode"]), s
        assert np.array_equal(out["nsift"][s], o["nsift"]), s
        assert np.array_equal(out["trace"][s], o["trace"]), s
        got = mimf[E * mi * offs[s]: E * mi * offs[s + 1]].reshape(E, mi, n)
        assert np.max(np.abs(got - o["member_imfs"])) <= 1e-13 * max(1.0, np.max(np.abs(xs[s]))), s


def test_automatic_task_count_is_one_for_the_batch(eng):
    """Automatic members per task on a batch of uneven lengths: one count for
    every signal (so the queue holds the same number of tasks per signal),
    taken from the longest signal.  Tasks only regroup the ensemble sum: the
    members' IMFs and sift counts equal those of one member per task bit for
    bit, and the ensemble means equal them to re-association."""
    rng = np.random.default_rng(33)
    lens = rng.integers(100, 2000, size=200)
    xs = []
    for n in lens:
        t = np.arange(n)
        f1, f2 = rng.uniform(0.01, 0.2, size=2)
        xs.append(np.sin(2 * np.pi * f1 * t) + 0.5 * np.sin(2 * np.pi * f2 * t)
                  + 0.05 * rng.standard_normal(n))
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    sig = np.array([0.2 * np.std(v) for v in xs])
    x = np.concatenate(xs)
    E, mi = 32, 4
    print("case: 200 uneven signals E=32 fp64, automatic vs one member per task", flush=True)
    auto = eng.run_batch(x, offs, E, sig, params(max_imf=mi, max_sifts=5, mpt=0, seed=3),
                         member_imfs=True)
    t_auto = eng.last_timing()
    one = eng.run_batch(x, offs, E, sig, params(max_imf=mi, max_sifts=5, mpt=1, seed=3),
                        member_imfs=True)
    nsig = len(lens)
    assert t_auto["ntasks"] % nsig == 0
    per_signal = t_auto["ntasks"] // nsig
    assert 1 < per_signal < E  # several members per task, the same for every length
    assert np.array_equal(auto["member_imfs"], one["member_imfs"])
    assert np.array_equal(auto["nmode"], one["nmode"])
    assert np.array_equal(auto["nsift"], one["nsift"])
    assert np.array_equal(auto["n_modes"], one["n_modes"])
    ma, m1 = means(auto, lens, mi), means(one, lens, mi)
    for s in range(nsig):
        scale = np.max(np.abs(xs[s]))
        assert np.max(np.abs(ma[s] - m1[s])) <= 1e-13 * scale, s


def test_one_wave_batch_with_sd_stop_matches_oracle(eng):
    """The SD stop (tau > 0: each IMF ends at the first sift whose SD is below
    tau, emd_ref.py:140-146) in a batched launch of one wave per member, with
    DSBT windows and mirror ends: picked members' extrema of every sift, sift
    and IMF counts and IMFs equal the oracle's (FP64, Philox noise)."""
    rng = np.random.default_rng(41)
    lens = rng.integers(500, 6000, size=150)
    xs = []
    for n in lens:
        t = np.arange(n)
        f1, f2, f3 = rng.uniform(0.002, 0.15, size=3)
        xs.append(np.sin(2 * np.pi * f1 * t) + 0.6 * np.sin(2 * np.pi * f2 * t + 0.3)
                  + 0.3 * np.sin(2 * np.pi * f3 * t) + 0.05 * rng.standard_normal(n))
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    sig = np.array([0.2 * np.std(v) for v in xs])
    E, mi, tau, ms = 16, 8, 0.2, 30
    p = params(max_imf=mi, max_sifts=ms, B=256, S=30, mpt=0, seed=9)
    p.tau = tau
    print("case: 150 signals E=16 fp64 tau=0.2 B=256 mirror, one wave per member", flush=True)
    out = eng.run_batch(np.concatenate(xs), offs, E, sig, p, member_imfs=True, trace=True)
    assert eng.last_timing()["threads"] == 64
    mimf = out["member_imfs"]
    order = np.argsort(lens)
    pick = order[np.linspace(0, len(xs) - 1, 6).astype(int)]
    for s in pick:
        n = int(lens[s])
        z = np.stack([orc.normals(9, int(s), e, n) for e in range(E)])
        o = orc.eemd(xs[s], z, sig[s], tau=tau, max_sifts=ms, max_imf=mi, end="mirror", B=256,
                     S=30, member_imfs=True, trace=True)
        assert np.array_equal(out["nmode"][s], o["nmode"]), s
        assert np.array_equal(out["nsift"][s], o["nsift"]), s
        assert np.array_equal(out["trace"][s], o["trace"]), s
        got = mimf[E * mi * offs[s]: E * mi * offs[s + 1]].reshape(E, mi, n)
        for e in range(E):
            for m in range(int(o["nmode"][e])):
                assert rel_l2(got[e, m], o["member_imfs"][e, m]) <= 1e-10, (s, e, m)
    # sift counts vary with the SD stop
    assert len(np.unique(out["nsift"])) > 1
