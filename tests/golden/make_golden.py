"""Write golden vectors by running the reference's own CPU path (emd_ref.py).

Run once in a container that has the reference at /root/reference:

    python tests/golden/make_golden.py            # EEMD / EMD / degenerate sets
    python tests/golden/make_golden.py ceemdan    # CEEMDAN set only

emd_ref.emd_1d's exact trajectory under numpy arithmetic (numpy's SIMD power
in the spline cube, pairwise np.sum in the SD test, searchsorted intervals)
is pinned by nothing but the reference itself, so the vectors come from
importing it.  Only inputs and outputs are stored (.npz, no pickles); the
reference's source never enters this repository.  The per-sift extrema digests
are recorded by wrapping emd_ref.envelope (the reference is not modified).
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(REF))
sys.path.insert(0, str(HERE.parents[1]))

import emd_ref as R  # noqa: E402  (the reference's CPU oracle)

from eemd4cuda_amd import signals  # noqa: E402

_M1 = np.uint64(0x9E3779B97F4A7C15)
_M2 = np.uint64(0xBF58476D1CE4E5B9)
_M3 = np.uint64(0x94D049BB133111EB)


def mix64(v):
    z = np.asarray(v, dtype=np.uint64) + _M1
    z = (z ^ (z >> np.uint64(30))) * _M2
    z = (z ^ (z >> np.uint64(27))) * _M3
    return z ^ (z >> np.uint64(31))


def digest(k):
    with np.errstate(over="ignore"):
        return int(np.sum(mix64(np.asarray(k, dtype=np.int64).astype(np.uint64)),
                          dtype=np.uint64)) if len(k) else 0


class Tracer:
    """Records (count, digest) of the knots of every envelope() call."""

    def __init__(self):
        self.calls = []
        self._orig = R.envelope

    def __enter__(self):
        def wrapped(t, x, want, end="natural"):
            if want == "max":
                k = np.where((x[1:-1] > x[:-2]) & (x[1:-1] > x[2:]))[0] + 1
            else:
                k = np.where((x[1:-1] < x[:-2]) & (x[1:-1] < x[2:]))[0] + 1
            self.calls.append((want, len(k), digest(k)))
            return self._orig(t, x, want, end)
        R.envelope = wrapped
        return self

    def __exit__(self, *a):
        R.envelope = self._orig

    def per_sift(self, per_mode, max_imf, max_sifts):
        out = np.zeros((max_imf, max_sifts, 4), dtype=np.uint64)
        it = iter(self.calls)
        for m, ns in enumerate(per_mode):
            for s in range(ns):
                a = next(it)
                b = next(it)
                assert a[0] == "max" and b[0] == "min"
                out[m, s] = (a[1], b[1], a[2], b[2])
        return out


def member_runs(x, noises, max_imf, **kw):
    E, N = noises.shape
    avg = np.zeros((max_imf, N))
    cnt = np.zeros(max_imf)
    nmode = np.zeros(E, np.int32)
    nsift = np.zeros(E, np.int32)
    per_mode = np.zeros((E, max_imf), np.int32)
    trace = np.zeros((E, max_imf, kw["max_sifts"], 4), np.uint64)
    imfs_all = []
    for e in range(E):
        with Tracer() as tr:
            imfs, resid, st = R.emd_1d(x + noises[e], max_imf=max_imf, **kw)
        nmode[e] = st["n_modes"]
        nsift[e] = st["total_sifts"]
        per_mode[e, :st["n_modes"]] = st["per_mode_sifts"]
        trace[e] = tr.per_sift(st["per_mode_sifts"], max_imf, kw["max_sifts"])
        imfs_all.append(imfs)
        for i, imf in enumerate(imfs):
            avg[i] += imf
            cnt[i] += 1
    maxm = int(nmode.max()) if E else 0
    avg = avg[:maxm] / np.where(cnt[:maxm] > 0, cnt[:maxm], 1)[:, None]
    resid = x - avg.sum(axis=0)
    return dict(avg=avg, resid=resid, nmode=nmode, nsift=nsift, per_mode=per_mode,
                trace=trace), imfs_all


def main():
    meta = dict(numpy=np.__version__, python=sys.version.split()[0])
    try:
        from numpy._core._multiarray_umath import __cpu_features__ as F
        meta["cpu_features"] = sorted(k for k, v in F.items() if v)
    except Exception:
        pass

    # ---- three-tone reference case: N=4096, E=64, 8 IMFs x 10 sifts, mirror ----
    x = signals.three_tone(4096)
    z = signals.standard_noise(64, 4096, seed=0)
    sigma = 0.2 * np.std(x)
    noises = z * sigma                      # emd_ref.py:204 rounding
    res, imfs_all = member_runs(x, noises, 8, tau=0.0, max_sifts=10, min_extrema=4,
                                end="mirror")
    mem = np.zeros((4, 8, 4096))
    for e in range(4):
        for i, imf in enumerate(imfs_all[e]):
            mem[e, i] = imf
    np.savez_compressed(HERE / "golden_three_tone.npz", x=x, z=z, sigma=sigma,
                        member_imfs_0_3=mem, **res)
    print("three_tone: n_modes=%d nmode=%s" % (res["avg"].shape[0], np.bincount(res["nmode"])))

    # ---- natural-end ensemble on the 2048-sample chirp ----
    xc = signals.chirp2048()
    zc = signals.standard_noise(8, 2048, seed=5)
    sc = 0.2 * np.std(xc)
    resn, imfs_n = member_runs(xc, zc * sc, 6, tau=0.0, max_sifts=10, min_extrema=4,
                               end="natural")
    memn = np.zeros((8, 6, 2048))
    for e in range(8):
        for i, imf in enumerate(imfs_n[e]):
            memn[e, i] = imf
    np.savez_compressed(HERE / "golden_chirp_natural.npz", x=xc, z=zc, sigma=sc,
                        member_imfs=memn, **resn)

    # ---- SD stopping (tau > 0), variable sifts, both ends, E=1 ----
    out = {}
    for end, tau in (("natural", 0.25), ("mirror", 0.2)):
        with Tracer() as tr:
            imfs, resid, st = R.emd_1d(xc, tau=tau, max_sifts=50, max_imf=-1,
                                       min_extrema=4, end=end)
        k = "%s_tau" % end
        out[k + "_imfs"] = np.array(imfs)
        out[k + "_resid"] = resid
        out[k + "_per_mode"] = np.array(st["per_mode_sifts"], np.int32)
        out[k + "_trace"] = tr.per_sift(st["per_mode_sifts"], len(imfs), 50)
    # envelopes of the chirp (single sift operator, test_emd_gpu.py:89-102)
    t = np.arange(2048, dtype=float)
    for end in ("natural", "mirror"):
        for want in ("max", "min"):
            env, p = R.envelope(t, xc, want, end)
            out["env_%s_%s" % (want, end)] = env
    np.savez_compressed(HERE / "golden_chirp_emd.npz", x=xc, **out)

    # ---- degenerate inputs ----
    cases = {}
    rng = np.random.default_rng(11)
    cases["constant"] = np.ones(64)
    cases["monotone"] = np.linspace(-3, 3, 64)
    plateau = rng.standard_normal(200)
    plateau[50:55] = 9.0
    plateau[120:130] = plateau[120]
    cases["plateau"] = plateau
    w = np.zeros(40)
    w[5] = -1; w[15] = -1; w[25] = -1; w[35] = -1        # minima only (flat tops)
    cases["minima_only"] = w
    cases["two_knots"] = np.sin(np.linspace(0, 3 * np.pi, 48))
    cases["short8"] = np.array([0.0, 1.0, -1.0, 2.0, -2.0, 1.5, -0.5, 0.25])
    cases["alternating"] = (np.arange(64) % 2) * 2.0 - 1.0 + 0.01 * np.arange(64)
    deg = {}
    for name, xx in cases.items():
        for end in ("natural", "mirror"):
            imfs, resid, st = R.emd_1d(xx, tau=0.0, max_sifts=10, max_imf=6,
                                       min_extrema=4, end=end)
            k = "%s_%s" % (name, end)
            deg[k + "_x"] = xx
            deg[k + "_imfs"] = np.array(imfs).reshape(len(imfs), len(xx))
            deg[k + "_resid"] = resid
            deg[k + "_per_mode"] = np.array(st["per_mode_sifts"], np.int32)
    np.savez_compressed(HERE / "golden_degenerate.npz", **deg)

    (HERE / "golden_meta.json").write_text(json.dumps(meta, indent=1))
    print("wrote", sorted(p.name for p in HERE.glob("golden_*")))


def golden_ceemdan():
    """emd_ref.ceemdan (emd_ref.py:220-259), serial, with the per-stage
    first-IMF sift counts and the noise-EMD IMF counts recorded by wrapping
    the reference's trial workers (the reference is not modified)."""
    out = {}
    x = signals.chirp2048()[:1024]
    orig_1imf, orig_eemd = R._trial_1imf, R._trial_eemd
    for end, tau, max_sifts in (("natural", 0.25, 50), ("mirror", 0.2, 20)):
        stage_ns, noise_n = [], []

        def t1(args):
            r = orig_1imf(args)
            stage_ns.append(r[1])
            return r

        def te(args):
            r = orig_eemd(args)
            noise_n.append(len(r[0]))
            return r

        R._trial_1imf, R._trial_eemd = t1, te
        try:
            c, resid, info = R.ceemdan(x, E=4, eps=0.2, parallel=False, seed=0, tau=tau,
                                       max_sifts=max_sifts, end=end)
        finally:
            R._trial_1imf, R._trial_eemd = orig_1imf, orig_eemd
        k = "%s_" % end
        out[k + "imfs"] = np.asarray(c)
        out[k + "resid"] = resid
        out[k + "stage_sifts"] = np.array(stage_ns, np.int32).reshape(-1, 4)
        out[k + "noise_nimf"] = np.array(noise_n, np.int32)
        out[k + "tau"] = tau
        out[k + "max_sifts"] = max_sifts
        print("ceemdan %s: %d IMFs, noise IMFs %s" % (end, len(c), noise_n))
    np.savez_compressed(HERE / "golden_ceemdan.npz", x=x, E=4, eps=0.2, seed=0, **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["ceemdan"]:
        golden_ceemdan()
        sys.exit(0)
    main()
