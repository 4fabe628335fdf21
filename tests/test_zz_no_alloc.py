"""GPU: the allocation-free entry point eemd_run_batch_dev_noalloc against
eemd_run_batch_dev, through the C ABI, on two calls in a row (the second
with a smaller batch), and its rejection of a batch that does not fit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(fp64, mpt, max_imf=6, B=256):
    from eemd4cuda_amd import _lib
    p = _lib.default_params()
    p.fp64 = fp64
    p.max_imf = max_imf
    p.block = B
    p.halo = 30 if fp64 else 14
    p.members_per_task = mpt
    p.seed = 3
    return p


def _batch(seed, lens):
    rng = np.random.default_rng(seed)
    xs = [np.sin(np.arange(n) * rng.uniform(0.01, 0.2)) + 0.3 * rng.standard_normal(n)
          for n in lens]
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    sig = np.array([0.2 * np.std(v) for v in xs])
    return np.concatenate(xs), offs, sig


def _run(eng, x, offs, E, sig, p, mi, noalloc):
    import torch
    dev = torch.device("cuda", 0)
    n, nsig = len(x), len(offs) - 1
    xd = torch.from_numpy(x).to(dev)
    avg = torch.full((mi * n,), 3.0, dtype=torch.float64, device=dev)
    res = torch.zeros(n, dtype=torch.float64, device=dev)
    nm = torch.zeros(nsig, dtype=torch.int32, device=dev)
    nmode = torch.zeros(nsig * E, dtype=torch.int32, device=dev)
    nsift = torch.zeros(nsig * E, dtype=torch.int32, device=dev)
    pm = torch.zeros(nsig * E * mi, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    eng.run_batch_dev(xd.data_ptr(), offs, E, sig, p, avg.data_ptr(), res.data_ptr(),
                      nm.data_ptr(), nmode.data_ptr(), nsift.data_ptr(),
                      per_mode_ptr=pm.data_ptr(), noalloc=noalloc)
    torch.cuda.synchronize()
    return dict(avg=avg.cpu().numpy(), resid=res.cpu().numpy(), n_modes=nm.cpu().numpy(),
                nmode=nmode.cpu().numpy(), nsift=nsift.cpu().numpy(), pm=pm.cpu().numpy())


@pytest.mark.parametrize("fp64", [1, 0])
def test_noalloc_equals_allocating_entry_point(fp64):
    from eemd4cuda_amd import _lib
    eng = _lib.engine()
    mi = 6
    for k, (lens, E) in enumerate((([9000, 20000, 4000], 6), ([3000, 5000], 3))):
        x, offs, sig = _batch(20 + k, lens)
        for mpt in (1, 0):
            print("case: noalloc fp64=%d call %d mpt=%d lens=%s E=%d" % (fp64, k, mpt, lens, E),
                  flush=True)
            a = _run(eng, x, offs, E, sig, _params(fp64, mpt), mi, False)
            b = _run(eng, x, offs, E, sig, _params(fp64, mpt), mi, True)
            for key in ("nmode", "nsift", "pm", "n_modes"):
                assert np.array_equal(a[key], b[key]), key
            if mpt == 1:
                # one member per task: the same member-order sums, bit for bit
                assert np.array_equal(a["avg"], b["avg"])
                assert np.array_equal(a["resid"], b["resid"])
            else:
                # the task split may differ (fewer workgroups fit): re-association only
                scale = np.max(np.abs(a["avg"]))
                assert np.max(np.abs(a["avg"] - b["avg"])) <= 1e-13 * scale
                assert np.max(np.abs(a["resid"] - b["resid"])) <= 1e-13 * scale


def test_noalloc_rejects_a_batch_that_does_not_fit():
    """Checked on the host before any device work: one member per task of a
    2^22-sample signal needs more accumulator memory than the scratch holds."""
    import torch
    from eemd4cuda_amd import _lib
    eng = _lib.engine()
    n, E = 1 << 22, 8
    dev = torch.device("cuda", 0)
    xd = torch.zeros(n, dtype=torch.float64, device=dev)
    out = torch.zeros(10 * n, dtype=torch.float64, device=dev)
    ints = torch.zeros(64, dtype=torch.int32, device=dev)
    p = _params(1, 1, max_imf=10)
    with pytest.raises(ValueError, match="allocation-free scratch"):
        eng.run_batch_dev(xd.data_ptr(), np.array([0, n]), E, np.array([0.1]), p,
                          out.data_ptr(), out.data_ptr(), ints.data_ptr(), ints.data_ptr(),
                          ints.data_ptr(), noalloc=True)
    p.members_per_task = 0
    p.flags = _lib.FLAG_GRAPH
    with pytest.raises(ValueError, match="FLAG_GRAPH"):
        eng.run_batch_dev(xd.data_ptr(), np.array([0, 5000]), 1, np.array([0.1]), p,
                          out.data_ptr(), out.data_ptr(), ints.data_ptr(), ints.data_ptr(),
                          ints.data_ptr(), noalloc=True)
