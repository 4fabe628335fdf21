"""CPU checks of the learned end extension's oracle and training data.

No reference counterpart exists. The oracle is pinned in five ways:
* the exponential and GELU against numpy's exp / tanh;
* the exact FP32 inference against a float64 forward pass;
* the float64 gradient against central finite differences;
* the virtual-extrema rule in closed form with an all-zero model, whose
  outputs are exactly 0;
* the resulting learned envelope against scipy's natural cubic spline
  through the extended knot set.
"""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import orc

import mlp_ref


def test_mlp_exp_and_gelu_against_numpy():
    x = np.linspace(-80, 80, 20001).astype(np.float32)
    e = orc.mlp_exp(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    assert np.max(np.abs(e / ref - 1)) < 3e-7
    g = orc.mlp_gelu(np.linspace(-12, 12, 4001).astype(np.float32)).astype(np.float64)
    xr = np.linspace(-12, 12, 4001).astype(np.float32).astype(np.float64)
    assert np.max(np.abs(g - mlp_ref.gelu(xr)) / np.maximum(1.0, np.abs(xr))) < 2e-6


def test_mlp_inference_matches_float64_forward():
    rng = np.random.default_rng(0)
    p = mlp_ref.random_params(rng)
    x = rng.standard_normal((40, 64)).astype(np.float32)
    y = orc.mlp_infer(p, x).astype(np.float64)
    # float64 forward with the same BF16 parameters, activations unrounded
    Ws, bs = mlp_ref.unpack(p, bf16=True)
    a = x.astype(np.float64)
    for l in range(4):
        z = a @ Ws[l] + bs[l]
        a = mlp_ref.gelu(z) if l < 3 else z
    assert np.max(np.abs(y - a)) <= 1e-4 * max(1.0, np.max(np.abs(a)))


def test_gradient_oracle_against_finite_differences():
    rng = np.random.default_rng(1)
    p = mlp_ref.random_params(rng).astype(np.float64)
    x = rng.standard_normal((8, 64))
    y = rng.standard_normal((8, 16))
    _, g = mlp_ref.loss_and_grad(p, x, y, bf16=False)
    idx = np.concatenate([rng.choice(len(p), 40, replace=False),
                          [0, 64 * 128 + 3, len(p) - 1, len(p) - 17]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp_ref.loss_and_grad(pp, x, y)[0] - mlp_ref.loss_and_grad(pm, x, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def _knots(h):
    kx = orc.extrema(h, True)
    kn = orc.extrema(h, False)
    return kx.astype(float), h[kx], kn.astype(float), h[kn]


def _signal(N=600, seed=2):
    rng = np.random.default_rng(seed)
    t = np.arange(N)
    return np.sin(2 * np.pi * t / 23.0) + 0.4 * np.sin(2 * np.pi * t / 7.3) + 0.05 * rng.standard_normal(N)


def test_zero_model_virtual_extrema_closed_form():
    h = _signal()
    N = len(h)
    tx, yx, tn, yn = _knots(h)
    zero = np.zeros(mlp_ref.offsets()[1], np.float32)
    vt, vy = orc.learned_virtual(zero, tx, yx, tn, yn, N)
    for end in (0, 1):
        x, (G, mu, sd) = orc.learned_features(tx, yx, tn, yn, N, end)
        # zero model -> outputs 0: first gap G - 1 from the end sample, then G
        for ev, (T, Y) in enumerate(((tx, yx), (tn, yn))):
            near = T[-1] if end else T[0]
            e_end = N - 1.0 if end else 0.0
            s = 1.0 if end else -1.0
            q = e_end + s * float(np.float32(np.float32(G) * np.float32(1.0) - np.float32(1.0)))
            if s * (q - near) < 2.0:
                q = near + 2.0 * s
            want = [q]
            for _ in range(3):
                q = q + s * float(max(np.float32(G), np.float32(2.0)))
                want.append(q)
            assert np.array_equal(vt[ev, end], want)
            assert np.all(vy[ev, end] == np.float64(np.float32(mu)))
        # features: first gap from the end sample, then gaps, as d / G - 1
        ls = (N - 1 - tx[::-1][:16]) if end else tx[:16]
        d = np.concatenate(([ls[0]], np.diff(ls)))
        assert np.allclose(x[:16], d / G - 1, atol=1e-6)


def test_learned_envelope_is_the_natural_spline_through_the_extended_knots():
    h = _signal()
    N = len(h)
    tx, yx, tn, yn = _knots(h)
    rng = np.random.default_rng(3)
    p = mlp_ref.random_params(rng, scale=0.3)
    orc.set_model(p)
    try:
        eu, el = orc.envelopes_learned(h)
    finally:
        orc.set_model(None)
    vt, vy = orc.learned_virtual(p, tx, yx, tn, yn, N)
    for ev, (T, Y, env) in enumerate(((tx, yx, eu), (tn, yn, el))):
        TT = np.concatenate((vt[ev, 0][::-1], T, vt[ev, 1]))
        YY = np.concatenate((vy[ev, 0][::-1], Y, vy[ev, 1]))
        assert np.all(np.diff(TT) >= 2.0)
        ref = CubicSpline(TT, YY, bc_type="natural")(np.arange(N))
        assert np.max(np.abs(env - ref)) <= 1e-9 * np.max(np.abs(YY))


def test_learned_emd_partition_of_unity_and_fallback():
    h = _signal(900, seed=4)
    rng = np.random.default_rng(5)
    orc.set_model(mlp_ref.random_params(rng, scale=0.3))
    try:
        o = orc.emd_1d(h, tau=0.0, max_sifts=10, max_imf=4, end="learned")
        # an envelope with < 16 extrema falls back to mirror: a short signal
        short = _signal(60, seed=6)
        a = orc.emd_1d(short, tau=0.0, max_sifts=10, max_imf=2, end="learned")
        b = orc.emd_1d(short, tau=0.0, max_sifts=10, max_imf=2, end="mirror")
    finally:
        orc.set_model(None)
    imfs, resid, st = o
    assert imfs.shape[0] == 4
    recon = imfs.sum(axis=0) + resid
    assert np.max(np.abs(recon - h)) <= 1e-13
    assert np.array_equal(a[0], b[0])


def test_training_examples_encode_the_long_signal_extrema():
    from eemd4cuda_amd import learned
    h = _signal(3000, seed=7)
    kx, kn = learned.strict_extrema(h)
    assert np.array_equal(kx, orc.extrema(h, True))
    assert np.array_equal(kn, orc.extrema(h, False))
    a, W = 700, 1200
    X, Y = learned._end_examples(h, kx, kn, [a], W)
    assert len(X) == 2
    e = a + W - 1
    # right end, maxima: decode the targets back to the true next extrema
    x, y = X[1], Y[1]
    inner = kx[(kx > a) & (kx < e)]
    outer = kx[kx >= e][:4]
    ls = e - inner[::-1][:16]
    ln_ = e - kn[(kn > a) & (kn < e)][::-1][:16]
    G = np.float32(np.float32(np.concatenate((np.diff(ls), np.diff(ln_))).astype(np.float32).sum())
                   / np.float32(30))
    d0 = np.exp(y[0]) * G - 1
    gaps = np.exp(y[1:4]) * G
    assert np.allclose(np.cumsum(np.concatenate(([e + d0], gaps))), outer, rtol=0, atol=1e-3)
    # the window's own extrema are the long signal's interior extrema
    kxw, _ = learned.strict_extrema(h[a:a + W])
    assert np.array_equal(kxw + a, inner)
    # features agree with the oracle's (same rule, different summation order)
    tx = kxw.astype(float)
    yx = h[a:a + W][kxw]
    _, knw = learned.strict_extrema(h[a:a + W])
    xo, _ = orc.learned_features(tx, yx, knw.astype(float), h[a:a + W][knw], W, 1)
    assert np.allclose(xo, x, atol=2e-5)
