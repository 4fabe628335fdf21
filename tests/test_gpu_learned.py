"""GPU: the learned end extension and its training kernels (FP64 parity, FP32 one sift).

* Inference: eemd_mlp_infer is bit-identical to the oracle's exact FP32
  inference.
* Training kernels (BF16 MFMA GEMMs, FP32 accumulation): the gradient agrees
  with the float64 numpy reference. That reference rounds the same operands
  to BF16 and is itself pinned by finite differences in the CPU tests. A
  short training run reduces the loss.
* In-kernel inference: FP64 EEMD with end='learned' reproduces the oracle,
  with bit-identical extrema traces and IMFs to rounding.
"""
import numpy as np
import pytest

from oracle import orc

import mlp_ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from eemd4cuda_amd import _lib
    return _lib.engine()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.device("cuda", 0))


def test_mlp_infer_bit_exact_with_oracle(eng):
    import torch
    from eemd4cuda_amd import learned
    rng = np.random.default_rng(0)
    p = mlp_ref.random_params(rng)
    x = rng.standard_normal((96, 64)).astype(np.float32) * 2
    print("case: mlp inference 96 samples", flush=True)
    net = learned.MLP(seed=1)
    net.set_params(p)
    xd = _dev(x)
    yd = torch.zeros((96, 16), dtype=torch.float32, device=xd.device)
    net.infer(xd.data_ptr(), 96, yd.data_ptr())
    y = yd.cpu().numpy()
    assert np.array_equal(y, orc.mlp_infer(p, x))


def test_mlp_gradient_matches_reference(eng):
    from eemd4cuda_amd import learned
    rng = np.random.default_rng(1)
    p = mlp_ref.random_params(rng)
    x = rng.standard_normal((64, 64)).astype(np.float32)
    y = rng.standard_normal((64, 16)).astype(np.float32)
    print("case: mlp gradient batch 64", flush=True)
    net = learned.MLP(seed=2)
    net.set_params(p)
    xd, yd = _dev(x), _dev(y)
    g, loss = net.grad(xd.data_ptr(), yd.data_ptr(), 64)
    l_ref, g_ref = mlp_ref.loss_and_grad(p, x, y, bf16=True)
    assert abs(loss - l_ref) <= 1e-4 * l_ref
    offs, _ = mlp_ref.offsets()
    for l, ((ow, ob), (i, j)) in enumerate(zip(offs, mlp_ref.SIZES)):
        gw, rw = g[ow:ow + i * j], g_ref[ow:ow + i * j]
        gb, rb = g[ob:ob + j], g_ref[ob:ob + j]
        ew = np.linalg.norm(gw - rw) / np.linalg.norm(rw)
        eb = np.linalg.norm(gb - rb) / np.linalg.norm(rb)
        print("layer", l, "rel err W %.2e b %.2e" % (ew, eb))
        # measured at most 3.1e-4 (weights) and 9.4e-5 (biases): FP32 accumulation of
        # BF16 products in MFMA order against the float64 reference of the same products
        assert ew <= 1e-3 and eb <= 1e-3, (l, ew, eb)
    # the parameters did not move
    assert np.array_equal(net.params(), p)


def test_mlp_training_reduces_loss(eng):
    from eemd4cuda_amd import learned
    rng = np.random.default_rng(2)
    teacher = mlp_ref.random_params(rng)
    x = rng.standard_normal((2048, 64)).astype(np.float32)
    y = mlp_ref.forward(teacher, x)[0][3].astype(np.float32)
    xd, yd = _dev(x), _dev(y)
    net = learned.MLP(seed=3)
    print("case: mlp training 300 steps", flush=True)
    first = None
    losses = []
    for s in range(300):
        b = (s % 16) * 128
        losses.append(net.step(xd[b:].data_ptr(), yd[b:].data_ptr(), 128, 1e-3))
        first = first if first is not None else losses[-1]
    last = float(np.mean(losses[-16:]))
    print("loss %.4f -> %.4f" % (first, last))
    assert np.all(np.isfinite(losses))
    assert last < 0.3 * first


@pytest.mark.parametrize("B,S,threads", [(0, 30, 0), (256, 30, 0), (256, 30, 64)])
def test_learned_ends_fp64_match_oracle(eng, B, S, threads):
    from eemd4cuda_amd import _lib, learned
    rng = np.random.default_rng(4)
    N = 6000
    t = np.arange(N)
    x = np.sin(2 * np.pi * t / 41.0) + 0.5 * np.sin(2 * np.pi * t / 9.7) + 0.2 * rng.standard_normal(N)
    z = rng.standard_normal((3, N))
    sigma = 0.2 * np.std(x)
    p = mlp_ref.random_params(rng, scale=0.3)
    print("case: learned ends fp64 N=6000 E=3 B=%d threads=%d" % (B, threads), flush=True)
    learned.install(p, eng)
    prm = _lib.default_params()
    prm.fp64 = 1
    prm.end_mode = _lib.END_MODES["learned"]
    prm.max_imf = 6
    prm.block = B
    prm.halo = S
    prm.members_per_task = 1
    prm.threads = threads
    out = eng.run_batch(x, np.array([0, N]), 3, np.array([sigma]), prm, noise=z,
                        member_imfs=True, trace=True)
    orc.set_model(p)
    try:
        o = orc.eemd(x, z, sigma, tau=0.0, max_sifts=10, max_imf=6, end="learned", B=B, S=S,
                     member_imfs=True, trace=True)
    finally:
        orc.set_model(None)
    assert np.array_equal(out["nmode"][0], o["nmode"])
    assert np.array_equal(out["trace"][0], o["trace"])
    mimf = out["member_imfs"].reshape(3, 6, N)
    d = np.max(np.abs(mimf - o["member_imfs"]))
    print("member IMF max|d| vs oracle:", d)
    assert d <= 1e-12
    # the learned ends changed the decomposition against mirror
    prm.end_mode = _lib.END_MODES["mirror"]
    mir = eng.run_batch(x, np.array([0, N]), 3, np.array([sigma]), prm, noise=z, member_imfs=True)
    assert not np.array_equal(mir["member_imfs"], out["member_imfs"])


def test_learned_without_model_is_rejected():
    from eemd4cuda_amd import _lib
    e2 = _lib.Engine(0)
    try:
        prm = _lib.default_params()
        prm.end_mode = _lib.END_MODES["learned"]
        with pytest.raises(ValueError, match="model"):
            e2.run_batch(np.sin(np.arange(500) * 0.1), np.array([0, 500]), 1, np.array([0.0]), prm)
    finally:
        e2.close()


def test_learned_ends_fp32_one_sift_matches_fp64(eng):
    """FP32 with learned ends.  For the first sift from x (no noise) both modes
    see the same knots, so the model input and its virtual extrema agree (the
    model always runs in FP32); the FP32 sift is then within FP32 rounding of
    the FP64 one, which is bit-exact with the oracle (test above)."""
    from eemd4cuda_amd import _lib, learned
    rng = np.random.default_rng(9)
    N = 20000
    t = np.arange(N)
    x = np.sin(2 * np.pi * t / 300.0) + 0.4 * np.sin(2 * np.pi * t / 23.0) + 0.05 * rng.standard_normal(N)
    p = mlp_ref.random_params(rng, scale=0.3)
    learned.install(p, eng)
    print("case: learned ends fp32 vs fp64 one sift N=20000", flush=True)
    outs = []
    for fp64, S in ((1, 30), (0, 14)):
        prm = _lib.default_params()
        prm.fp64 = fp64
        prm.end_mode = _lib.END_MODES["learned"]
        prm.max_imf = 1
        prm.max_sifts = 1
        prm.block = 1024
        prm.halo = S
        prm.members_per_task = 1
        outs.append(eng.run_batch(x, np.array([0, N]), 1, np.array([0.0]), prm,
                                  noise=np.zeros((1, N)), member_imfs=True))
    a = np.asarray(outs[0]["member_imfs"][:N], np.float64)
    b = np.asarray(outs[1]["member_imfs"][:N], np.float64)
    err = np.max(np.abs(a - b)) / np.max(np.abs(x))
    print("fp32 vs fp64 learned one-sift max rel err %.2e" % err)
    assert err <= 2e-5
    # and the learned ends differ from mirror ends near the ends
    prm.end_mode = _lib.END_MODES["mirror"]
    mir = eng.run_batch(x, np.array([0, N]), 1, np.array([0.0]), prm, noise=np.zeros((1, N)),
                        member_imfs=True)
    assert not np.array_equal(np.asarray(mir["member_imfs"][:N]), outs[1]["member_imfs"][:N])


def test_packaged_model_cuts_edge_error(eng):
    """The packaged model against mirror ends on 128 held-out windows (seed 5,
    never used in training): the long signal's FP64 decomposition judges the
    first / last 256 samples of each window's IMFs.  The aggregate error of
    IMFs 1-3 falls by at least 30%."""
    from eemd4cuda_amd import learned
    params = learned.load(learned.default_model_path())
    print("case: edge error, packaged model, 128 windows, E=1 fp64", flush=True)
    ev = learned.edge_error(params, n_windows=128, engine=eng)
    print("mirror", ["%.3f" % v for v in ev["mirror"]], "learned",
          ["%.3f" % v for v in ev["learned"]], "ratio", ["%.3f" % v for v in ev["ratio"]])
    assert all(r <= 0.7 for r in ev["ratio"][:3]), ev["ratio"]
    assert all(r <= 0.7 for r in ev["ratio_median_window"][:3]), ev["ratio_median_window"]


def test_packaged_model_cuts_edge_error_of_fp32_eemd(eng):
    """As above, with the windows run the way the at-scale batch runs them:
    FP32 EEMD with 16 members (noise 0.2 std), each member receiving the long
    signal's member noise restricted to the window."""
    from eemd4cuda_amd import learned
    params = learned.load(learned.default_model_path())
    print("case: edge error, packaged model, 64 windows, E=16 fp32", flush=True)
    ev = learned.edge_error(params, n_windows=64, E=16, eps=0.2, fp64=0, engine=eng)
    print("mirror", ["%.3f" % v for v in ev["mirror"]], "learned",
          ["%.3f" % v for v in ev["learned"]], "ratio", ["%.3f" % v for v in ev["ratio"]])
    assert all(r <= 0.7 for r in ev["ratio"][:3]), ev["ratio"]


def test_training_on_chirp_windows_reduces_validation_loss(eng):
    """A small run of the training recipe: windows cropped from 16,384-sample
    chirp/AM signals (seed 3), targets from the long signals' FP64 sift, three
    epochs of Adam on the GPU kernels; the held-out loss falls."""
    from eemd4cuda_amd import learned
    print("case: training smoke 400 windows 3 epochs", flush=True)
    ds = learned.make_dataset(n_windows=400, seed=3, engine=eng)
    assert ds.x.shape[1] == 64 and ds.y.shape[1] == 16 and len(ds.x) >= 2000
    params, hist = learned.train(ds, epochs=3, lr=1e-3, batch=256, seed=0)
    print("history", hist)
    assert np.all(np.isfinite(params))
    assert hist[-1][2] < hist[0][2]
    assert hist[-1][1] < hist[0][1]
