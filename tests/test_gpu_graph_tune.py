"""GPU: hipGraph replay of the launch sequence and the autotuner.

A replayed graph must give exactly the eager result, call after call, with
the program's own eager work (other decompositions, synchronisations) in
between.  A tuned launch configuration changes the schedule only: FP64
extrema / IMF counts stay identical and the ensemble mean moves by at most
re-association rounding.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from eemd4cuda_amd import _lib
    return _lib.engine()


def _params(fp64, B=256, S=30, max_imf=6):
    from eemd4cuda_amd import _lib
    p = _lib.default_params()
    p.fp64 = fp64
    p.max_imf = max_imf
    p.block = B
    p.halo = S if fp64 else 14
    p.seed = 5
    return p


def _batch(seed, lens):
    rng = np.random.default_rng(seed)
    xs = [np.sin(np.arange(n) * rng.uniform(0.01, 0.2)) + 0.3 * rng.standard_normal(n)
          for n in lens]
    offs = np.concatenate(([0], np.cumsum(lens))).astype(np.int64)
    sig = np.array([0.2 * np.std(v) for v in xs])
    return np.concatenate(xs), offs, sig


@pytest.mark.parametrize("fp64", [1, 0])
def test_graph_replays_equal_eager(eng, fp64):
    import torch
    x, offs, sig = _batch(1, [3000, 7000, 12000])
    xo, offo, sigo = _batch(2, [2500, 4000])
    E = 6
    print("case: graph fp64=%d" % fp64, flush=True)
    ref = eng.run_batch(x, offs, E, sig, _params(fp64))
    cap0, _ = eng.graph_stats()
    for rep in range(3):
        g = eng.run_batch(x, offs, E, sig, _params(fp64), graph=True)
        cap, _ = eng.graph_stats()
        assert cap == cap0 + 1  # captured once, replayed afterwards
        assert g["timing"]["mean_ms"] == 0.0  # served by the graph
        assert np.array_equal(g["avg"], ref["avg"]), rep
        assert np.array_equal(g["resid"], ref["resid"]), rep
        assert np.array_equal(g["nmode"], ref["nmode"]), rep
        assert np.array_equal(g["nsift"], ref["nsift"]), rep
        # the program's own eager work between replays (smaller batch: same buffers)
        o = eng.run_batch(xo, offo, E, sigo, _params(fp64))
        assert np.all(o["n_modes"] > 0)
        torch.cuda.synchronize()
    # a different batch through the graph path re-captures and is still exact
    o_ref = eng.run_batch(xo, offo, E, sigo, _params(fp64))
    o_g = eng.run_batch(xo, offo, E, sigo, _params(fp64), graph=True)
    assert np.array_equal(o_g["avg"], o_ref["avg"])


def test_graph_device_buffers_replay(eng):
    import torch
    from eemd4cuda_amd import _lib
    x, offs, sig = _batch(3, [5000, 9000])
    E, mi = 4, 6
    dev = torch.device("cuda", 0)
    xd = torch.from_numpy(x).to(dev)
    outs = []
    for graph in (False, True, True, True):
        avg = torch.full((mi * len(x),), 7.0, dtype=torch.float64, device=dev)
        res = torch.zeros(len(x), dtype=torch.float64, device=dev)
        nm = torch.zeros(2, dtype=torch.int32, device=dev)
        nmode = torch.zeros(2 * E, dtype=torch.int32, device=dev)
        nsift = torch.zeros(2 * E, dtype=torch.int32, device=dev)
        p = _params(1, max_imf=mi)
        p.flags = _lib.FLAG_GRAPH if graph else 0
        print("case: device graph", graph, flush=True)
        eng.run_batch_dev(xd.data_ptr(), offs, E, sig, p, avg.data_ptr(), res.data_ptr(),
                          nm.data_ptr(), nmode.data_ptr(), nsift.data_ptr())
        torch.cuda.synchronize()
        outs.append((avg.cpu().numpy(), res.cpu().numpy(), nmode.cpu().numpy()))
    # fresh output tensors each call: the graph is re-captured for new addresses
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)


def _dev_run(eng, xd, offs, E, sig, p, mi):
    import torch
    dev = xd.device
    n = xd.numel()
    nsig = len(offs) - 1
    avg = torch.zeros(mi * n, dtype=torch.float64, device=dev)
    res = torch.zeros(n, dtype=torch.float64, device=dev)
    nm = torch.zeros(nsig, dtype=torch.int32, device=dev)
    nmode = torch.zeros(nsig * E, dtype=torch.int32, device=dev)
    nsift = torch.zeros(nsig * E, dtype=torch.int32, device=dev)
    eng.run_batch_dev(xd.data_ptr(), offs, E, sig, p, avg.data_ptr(), res.data_ptr(),
                      nm.data_ptr(), nmode.data_ptr(), nsift.data_ptr())
    torch.cuda.synchronize()
    tm = eng.last_timing()
    return dict(avg=avg.cpu().numpy(), resid=res.cpu().numpy(), nmode=nmode.cpu().numpy(),
                nsift=nsift.cpu().numpy(), timing=tm)


@pytest.mark.parametrize("fp64", [1, 0])
def test_autotune_choice_is_applied_and_exact(eng, fp64):
    import torch
    x, offs, sig = _batch(4, [4000, 16000, 40000, 9000])
    E, mi = 12, 6
    xd = torch.from_numpy(x).to(torch.device("cuda", 0))
    p = _params(fp64, B=512, max_imf=mi)
    print("case: autotune fp64=%d" % fp64, flush=True)
    base = _dev_run(eng, xd, offs, E, sig, p, mi)
    best, ms = eng.autotune(xd.data_ptr(), offs, E, sig, p, reps=1)
    assert best.threads in (64, 256, 512, 1024)
    assert 0 <= best.members_per_task <= E  # 0: the automatic rule
    assert ms > 0
    if not fp64:
        assert best.wave_windows != 0
    # explicit best and the cached choice (all-automatic parameters) agree bit for bit
    exp = _dev_run(eng, xd, offs, E, sig, best, mi)
    cached = _dev_run(eng, xd, offs, E, sig, p, mi)
    assert cached["timing"]["threads"] == best.threads
    assert np.array_equal(exp["avg"], cached["avg"])
    assert np.array_equal(exp["nsift"], cached["nsift"])
    if fp64:
        # same members, same extrema: only the ensemble sum is re-associated
        assert np.array_equal(exp["nmode"], base["nmode"])
        assert np.array_equal(exp["nsift"], base["nsift"])
        scale = np.max(np.abs(base["avg"]))
        assert np.max(np.abs(exp["avg"] - base["avg"])) <= 1e-13 * scale


def test_graph_new_output_buffer_recaptures(eng):
    """Only the mean output buffer changes between two graph calls: the call
    is captured again and fills the new buffer (the old one is left alone)."""
    import torch
    from eemd4cuda_amd import _lib
    x, offs, sig = _batch(6, [6000, 3000])
    E, mi = 4, 6
    dev = torch.device("cuda", 0)
    xd = torch.from_numpy(x).to(dev)
    n = len(x)
    res = torch.zeros(n, dtype=torch.float64, device=dev)
    nm = torch.zeros(2, dtype=torch.int32, device=dev)
    nmode = torch.zeros(2 * E, dtype=torch.int32, device=dev)
    nsift = torch.zeros(2 * E, dtype=torch.int32, device=dev)
    p = _params(1, max_imf=mi)
    ref = _dev_run(eng, xd, offs, E, sig, p, mi)
    avgs = []
    for k in range(2):
        avg = torch.full((mi * n,), -1.0, dtype=torch.float64, device=dev)
        cap0, _ = eng.graph_stats()
        p.flags = _lib.FLAG_GRAPH
        print("case: graph with output buffer", k, flush=True)
        eng.run_batch_dev(xd.data_ptr(), offs, E, sig, p, avg.data_ptr(), res.data_ptr(),
                          nm.data_ptr(), nmode.data_ptr(), nsift.data_ptr())
        torch.cuda.synchronize()
        cap1, _ = eng.graph_stats()
        assert cap1 == cap0 + 1
        avgs.append(avg)
    assert np.array_equal(avgs[1].cpu().numpy(), ref["avg"])
    # the first buffer holds the first call's output and was not written again
    avgs[0].fill_(-1.0)
    torch.cuda.synchronize()
    p.flags = _lib.FLAG_GRAPH
    eng.run_batch_dev(xd.data_ptr(), offs, E, sig, p, avgs[1].data_ptr(), res.data_ptr(),
                      nm.data_ptr(), nmode.data_ptr(), nsift.data_ptr())
    torch.cuda.synchronize()
    assert np.all(avgs[0].cpu().numpy() == -1.0)
    assert np.array_equal(avgs[1].cpu().numpy(), ref["avg"])


def test_graph_replay_after_staging_buffers_grow():
    """Host-buffer path on a fresh context: an eager call with more IMFs grows
    (frees and reallocates) the context's output staging between two graph
    calls of the same batch.  The second graph call must write the staging
    buffers the context holds now (a replay into the freed one would return
    the eager call's data)."""
    from eemd4cuda_amd import _lib
    e2 = _lib.Engine(0)
    try:
        x, offs, sig = _batch(7, [4000, 5000])
        E = 4
        ref = e2.run_batch(x, offs, E, sig, _params(1, max_imf=3))
        g1 = e2.run_batch(x, offs, E, sig, _params(1, max_imf=3), graph=True)
        big = e2.run_batch(x, offs, E, sig, _params(1, max_imf=12))
        assert np.all(big["n_modes"] > 3)
        print("case: graph after staging growth", flush=True)
        g2 = e2.run_batch(x, offs, E, sig, _params(1, max_imf=3), graph=True)
        for g in (g1, g2):
            assert np.array_equal(g["avg"], ref["avg"])
            assert np.array_equal(g["resid"], ref["resid"])
            assert np.array_equal(g["nsift"], ref["nsift"])
            assert np.array_equal(g["n_modes"], ref["n_modes"])
    finally:
        e2.close()


def test_async_calls_on_two_streams_are_ordered(eng):
    """Two asynchronous calls of one context on different streams share the
    context's scratch: the second waits for the first, both results exact."""
    import torch
    from eemd4cuda_amd import _lib
    dev = torch.device("cuda", 0)
    jobs = [_batch(8, [20000, 15000]), _batch(9, [9000, 30000, 4000])]
    E, mi = 6, 6
    refs = [eng.run_batch(x, o, E, s, _params(1, max_imf=mi)) for x, o, s in jobs]
    streams = [torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)]
    bufs = []
    for (x, offs, sig), st in zip(jobs, streams):
        n, nsig = len(x), len(offs) - 1
        b = dict(x=torch.from_numpy(x).to(dev),
                 avg=torch.zeros(mi * n, dtype=torch.float64, device=dev),
                 res=torch.zeros(n, dtype=torch.float64, device=dev),
                 nm=torch.zeros(nsig, dtype=torch.int32, device=dev),
                 nmode=torch.zeros(nsig * E, dtype=torch.int32, device=dev),
                 nsift=torch.zeros(nsig * E, dtype=torch.int32, device=dev))
        bufs.append(b)
    torch.cuda.synchronize()
    print("case: two async calls on two streams", flush=True)
    for (x, offs, sig), st, b in zip(jobs, streams, bufs):
        p = _params(1, max_imf=mi)
        eng.run_batch_dev(b["x"].data_ptr(), offs, E, sig, p, b["avg"].data_ptr(),
                          b["res"].data_ptr(), b["nm"].data_ptr(), b["nmode"].data_ptr(),
                          b["nsift"].data_ptr(), stream=st.cuda_stream)
    torch.cuda.synchronize()
    for ref, b in zip(refs, bufs):
        assert np.array_equal(b["avg"].cpu().numpy(), ref["avg"])
        assert np.array_equal(b["nsift"].cpu().numpy(), ref["nsift"].ravel())


def test_autotune_fp64_learned_ends(eng):
    """FP64 with learned ends: the tuner skips the workgroup size that cannot
    hold the model in LDS and returns a valid choice."""
    import torch
    from eemd4cuda_amd import learned
    x, offs, sig = _batch(10, [6000, 12000])
    E, mi = 4, 4
    learned.install(learned.load(learned.default_model_path()), eng)
    xd = torch.from_numpy(x).to(torch.device("cuda", 0))
    p = _params(1, B=512, max_imf=mi)
    from eemd4cuda_amd import _lib
    p.end_mode = _lib.END_MODES["learned"]
    print("case: autotune fp64 learned", flush=True)
    best, ms = eng.autotune(xd.data_ptr(), offs, E, sig, p, reps=1)
    assert best.threads in (256, 512) and ms > 0
    p.end_mode = _lib.END_MODES["mirror"]


def test_autotune_block_only_with_flag(eng):
    """The tuner changes the block size only under EEMD_FLAG_TUNE_BLOCK, and a
    later automatic call uses the tuned block only when it sets the flag too:
    its result then equals an explicit call with that block."""
    import torch
    from eemd4cuda_amd import _lib
    x, offs, sig = _batch(11, [6000, 20000, 3000])
    E, mi = 10, 5
    xd = torch.from_numpy(x).to(torch.device("cuda", 0))
    p = _params(0, B=1024, max_imf=mi)
    print("case: autotune without the block flag", flush=True)
    best, _ = eng.autotune(xd.data_ptr(), offs, E, sig, p, reps=1)
    assert best.block == 1024
    p.flags = _lib.FLAG_TUNE_BLOCK
    print("case: autotune with the block flag", flush=True)
    best, _ = eng.autotune(xd.data_ptr(), offs, E, sig, p, reps=1)
    p.flags = 0
    assert best.block in (128, 256, 512, 1024, 2048)
    exp = _params(0, B=best.block, max_imf=mi)
    exp.threads, exp.members_per_task, exp.wave_windows = (best.threads, best.members_per_task,
                                                           best.wave_windows)
    a = _dev_run(eng, xd, offs, E, sig, exp, mi)
    auto = _params(0, B=1024, max_imf=mi)
    auto.flags = _lib.FLAG_TUNE_BLOCK
    b = _dev_run(eng, xd, offs, E, sig, auto, mi)
    assert np.array_equal(a["avg"], b["avg"]) and np.array_equal(a["nsift"], b["nsift"])
    # without the flag the automatic call keeps the requested block
    c = _dev_run(eng, xd, offs, E, sig, _params(0, B=1024, max_imf=mi), mi)
    ref = _params(0, B=1024, max_imf=mi)
    ref.threads, ref.members_per_task, ref.wave_windows = (best.threads, best.members_per_task,
                                                           best.wave_windows)
    d = _dev_run(eng, xd, offs, E, sig, ref, mi)
    assert np.array_equal(c["avg"], d["avg"])
