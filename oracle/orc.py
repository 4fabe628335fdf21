"""ctypes wrapper over the C oracle (``oracle/_build/liborc.so``).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.  Every function cites
the reference code it restates (file:line relative to the reference root).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "liborc.so"
_lib = None

END = {"natural": 0, "mirror": 1, "learned": 2}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_model_keep = None


def build(force: bool = False) -> Path:
    """Compile the oracle with ``oracle/Makefile`` (gcc, -ffp-contract=off)."""
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-B" if force else "-s", "-C", str(_HERE)],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = C.CDLL(str(_LIB_PATH))
        L.orc_extrema.restype = C.c_int64
        L.orc_extrema.argtypes = [_dp, C.c_int64, C.c_int, C.c_void_p]
        L.orc_pairwise_sum.restype = C.c_double
        L.orc_pairwise_sum.argtypes = [_dp, C.c_int64]
        L.orc_dsbt.restype = None
        L.orc_dsbt.argtypes = [_dp, _dp, _dp, _dp, C.c_int64, C.c_int64, C.c_int64, _dp]
        L.orc_envelope.restype = C.c_int64
        L.orc_envelope.argtypes = [_dp, C.c_int64, C.c_int, C.c_int, C.c_int64, C.c_int64, _dp]
        L.orc_spline_knots.restype = C.c_int64
        L.orc_spline_knots.argtypes = [_dp, C.c_int64, C.c_int, C.c_int, C.c_int64,
                                       C.c_int64, _dp, _dp, _dp]
        L.orc_emd_1d.restype = C.c_int
        L.orc_emd_1d.argtypes = [_dp, C.c_int64, C.c_double, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.c_int64, C.c_int64, C.c_int, _dp, _dp, _ip,
                                 C.POINTER(C.c_int), C.c_void_p]
        L.orc_eemd.restype = C.c_int
        L.orc_eemd.argtypes = [_dp, C.c_int64, C.c_int, _dp, C.c_double, C.c_double,
                               C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64,
                               C.c_int, C.c_int, _dp, _dp, _ip, _ip, C.c_void_p,
                               C.c_void_p, C.c_void_p]
        L.orc_philox4x32_10.restype = None
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_normals.restype = None
        L.orc_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, _dp]
        L.orc_mix64.restype = C.c_uint64
        L.orc_mix64.argtypes = [C.c_uint64]
        L.orc_set_model.restype = None
        L.orc_set_model.argtypes = [C.c_void_p]
        L.orc_mlp_nparam.restype = C.c_int
        for nm, a in (("orc_mlp_exp_n", _fp), ("orc_mlp_gelu_n", _fp)):
            getattr(L, nm).restype = None
            getattr(L, nm).argtypes = [a, _fp, C.c_int64]
        L.orc_mlp_infer_n.restype = None
        L.orc_mlp_infer_n.argtypes = [_u16p, _fp, _fp, C.c_int64]
        L.orc_f32_to_bf16_n.restype = None
        L.orc_f32_to_bf16_n.argtypes = [_fp, _u16p, C.c_int64]
        L.orc_learned_features.restype = None
        L.orc_learned_features.argtypes = [_dp, _dp, C.c_int64, _dp, _dp, C.c_int64, C.c_int64,
                                           C.c_int, _fp, _dp, _fp]
        L.orc_learned_virtual.restype = None
        L.orc_learned_virtual.argtypes = [_dp, _dp, C.c_int64, _dp, _dp, C.c_int64, C.c_int64,
                                          _u16p, _dp, _dp]
        L.orc_envelopes_learned_pub.restype = C.c_int
        L.orc_envelopes_learned_pub.argtypes = [_dp, C.c_int64, C.c_int64, C.c_int64, _dp, _dp]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def extrema(h, want_max: bool) -> np.ndarray:
    """Strict local extrema indices (emd_ref.py:84-87)."""
    h = _f64(h)
    k = np.zeros(max(len(h), 1), dtype=np.int64)
    p = lib().orc_extrema(h, len(h), int(bool(want_max)), k.ctypes.data)
    return k[:p].copy()


def pairwise_sum(a) -> float:
    """numpy's pairwise summation (the np.sum of emd_ref.py:141-142)."""
    a = _f64(a)
    return float(lib().orc_pairwise_sum(a, len(a)))


def dsbt(d, l, u, r, B: int = 0, S: int = 0) -> np.ndarray:
    """Decay-seamed block Thomas (emd_gpu_v5_dsbt.py:51-74, 223-237);
    B<=0 -> exact Thomas (emd_ref.py:50-65)."""
    d, l, u, r = map(_f64, (d, l, u, r))
    x = np.zeros(len(d))
    lib().orc_dsbt(d, l, u, r, len(d), int(B), int(S), x)
    return x


def envelope(h, want_max: bool, end: str = "natural", B: int = 0, S: int = 0):
    """Cubic-spline envelope on t=0..N-1 (emd_ref.py:77-103).  Returns (env, p)."""
    h = _f64(h)
    env = np.zeros(len(h))
    p = lib().orc_envelope(h, len(h), int(bool(want_max)), END[end], int(B), int(S), env)
    return env, int(p)


def spline_knots(h, want_max: bool, end: str = "natural", B: int = 0, S: int = 0):
    """Extended knots (T, Y) and second derivatives M of the envelope spline."""
    h = _f64(h)
    n = len(h) + 4
    T = np.zeros(n); Y = np.zeros(n); M = np.zeros(n)
    P = lib().orc_spline_knots(h, len(h), int(bool(want_max)), END[end], int(B), int(S), T, Y, M)
    return T[:P].copy(), Y[:P].copy(), M[:P].copy()


def emd_1d(x, tau=0.25, max_sifts=50, max_imf=-1, min_extrema=4, end="natural",
           B: int = 0, S: int = 0, max_imf_cap: int = 64, trace: bool = False):
    """Single-signal EMD (emd_ref.py:108-156).

    Returns (imfs [n_modes, N], resid, stats) with stats like emd_ref's plus
    'trace' ([n_modes, max_sifts, 4] uint64: nmax, nmin, hash_max, hash_min
    of the knots each executed sift used) when requested."""
    x = _f64(x)
    N = len(x)
    cap = max_imf_cap if max_imf == -1 else min(max_imf, max_imf_cap)
    cap = max(cap, 1)
    imfs = np.zeros((cap, N))
    resid = np.zeros(N)
    pm = np.zeros(cap, dtype=np.int32)
    ts = C.c_int(0)
    tr = np.zeros((cap, max(max_sifts, 1), 4), dtype=np.uint64) if trace else None
    nm = lib().orc_emd_1d(x, N, float(tau), int(max_sifts), int(max_imf), int(min_extrema),
                          END[end], int(B), int(S), cap, imfs, resid, pm, C.byref(ts),
                          tr.ctypes.data if trace else None)
    if nm < 0:
        raise MemoryError("oracle allocation failed")
    stats = dict(n_modes=nm, total_sifts=int(ts.value), per_mode_sifts=[int(v) for v in pm[:nm]])
    if trace:
        stats["trace"] = tr[:nm]
    return imfs[:nm].copy(), resid, stats


def eemd(x, z, sigma, tau=0.0, max_sifts=10, max_imf=8, min_extrema=4, end="mirror",
         B: int = 0, S: int = 0, nthreads: int = 0, member_imfs: bool = False,
         trace: bool = False):
    """EEMD with an explicit standard-normal matrix z [E, N] (member e
    decomposes x + z[e]*sigma) and max_imf honoured per member; level-wise
    sequential mean (emd_ref.py:199-217).  max_imf must be >= 1.

    Returns dict(avg [n_modes, N], resid [N], n_modes, nmode [E], nsift [E],
    per_mode [E, max_imf], member_imfs [E, max_imf, N] opt, trace opt)."""
    x = _f64(x)
    z = _f64(z)
    E, N = z.shape
    assert N == len(x) and max_imf >= 1
    cap = int(max_imf)
    avg = np.zeros((cap, N))
    resid = np.zeros(N)
    nmode = np.zeros(E, dtype=np.int32)
    nsift = np.zeros(E, dtype=np.int32)
    pm = np.zeros((E, cap), dtype=np.int32)
    mi = np.zeros((E, cap, N)) if member_imfs else None
    tr = np.zeros((E, cap, max(max_sifts, 1), 4), dtype=np.uint64) if trace else None
    nm = lib().orc_eemd(x, N, E, z, float(sigma), float(tau), int(max_sifts), int(max_imf),
                        int(min_extrema), END[end], int(B), int(S), cap, int(nthreads), avg,
                        resid, nmode, nsift, pm.ctypes.data,
                        mi.ctypes.data if member_imfs else None,
                        tr.ctypes.data if trace else None)
    if nm < 0:
        raise MemoryError("oracle allocation failed")
    out = dict(avg=avg[:nm].copy(), resid=resid, n_modes=int(nm), nmode=nmode, nsift=nsift,
               per_mode=pm)
    if member_imfs:
        out["member_imfs"] = mi
    if trace:
        out["trace"] = tr
    return out


def philox4x32_10(ctr, key) -> np.ndarray:
    """Philox4x32-10 block function (Salmon et al., SC'11)."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(c, k, o)
    return o


def normals(seed: int, signal: int, member: int, N: int) -> np.ndarray:
    """Counter-based standard normals the GPU path generates in-kernel."""
    out = np.zeros(N)
    lib().orc_normals(int(seed) & 0xFFFFFFFFFFFFFFFF, int(signal), int(member), int(N), out)
    return out


def normals_matrix(seed: int, signal: int, E: int, N: int) -> np.ndarray:
    return np.stack([normals(seed, signal, e, N) for e in range(E)]) if E else np.zeros((0, N))


def mix64(v: int) -> int:
    return int(lib().orc_mix64(int(v) & 0xFFFFFFFFFFFFFFFF))


def hash_indices(k) -> int:
    """Order-independent digest of an extrema index set (sum of splitmix64)."""
    s = 0
    for v in np.asarray(k, dtype=np.int64):
        s = (s + mix64(int(v))) & 0xFFFFFFFFFFFFFFFF
    return s


# ---------------------------------------------------------------------------
# Learned end extension (the engine's end mode "learned"; no reference
# counterpart): exact inference arithmetic and virtual-extrema rule.
# ---------------------------------------------------------------------------
def bf16(params) -> np.ndarray:
    """FP32 parameters rounded to BF16 (round to nearest even), as uint16."""
    a = np.ascontiguousarray(params, dtype=np.float32)
    out = np.zeros(a.size, dtype=np.uint16)
    lib().orc_f32_to_bf16_n(a.ravel(), out, a.size)
    return out


def set_model(params):
    """Model used by end='learned' (FP32 parameters; None: mirror only)."""
    global _model_keep
    if params is None:
        _model_keep = None
        lib().orc_set_model(None)
        return
    _model_keep = bf16(params)
    if _model_keep.size != lib().orc_mlp_nparam():
        raise ValueError("wrong parameter count")
    lib().orc_set_model(_model_keep.ctypes.data)


def mlp_exp(x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros_like(x)
    lib().orc_mlp_exp_n(x.ravel(), y.ravel(), x.size)
    return y


def mlp_gelu(x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros_like(x)
    lib().orc_mlp_gelu_n(x.ravel(), y.ravel(), x.size)
    return y


def mlp_infer(params, x) -> np.ndarray:
    """Exact FP32 inference over BF16 parameters: x [n, 64] -> [n, 16]."""
    w = bf16(params)
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 64)
    y = np.zeros((x.shape[0], 16), dtype=np.float32)
    lib().orc_mlp_infer_n(w, x, y, x.shape[0])
    return y


def learned_features(tx, yx, tn, yn, N, end):
    """Model input of one end (0 left, 1 right) and its (G, mu, sd) scales."""
    tx, yx, tn, yn = map(_f64, (tx, yx, tn, yn))
    x = np.zeros(64, dtype=np.float32)
    ls0 = np.zeros(2)
    sc = np.zeros(3, dtype=np.float32)
    lib().orc_learned_features(tx, yx, len(tx), tn, yn, len(tn), int(N), int(end), x, ls0, sc)
    return x, sc


def learned_virtual(params, tx, yx, tn, yn, N):
    """Virtual extrema [env][end][k] positions and values (k=0 nearest)."""
    tx, yx, tn, yn = map(_f64, (tx, yx, tn, yn))
    vt = np.zeros(16)
    vy = np.zeros(16)
    lib().orc_learned_virtual(tx, yx, len(tx), tn, yn, len(tn), int(N), bf16(params), vt, vy)
    return vt.reshape(2, 2, 4), vy.reshape(2, 2, 4)


def envelopes_learned(h, B=0, S=0):
    """Upper and lower envelopes with learned ends (model from set_model)."""
    h = _f64(h)
    eu = np.zeros(len(h))
    el = np.zeros(len(h))
    if lib().orc_envelopes_learned_pub(h, len(h), int(B), int(S), eu, el) != 0:
        raise MemoryError
    return eu, el


def ceemdan(x, W, eps=0.2, tau=0.25, max_sifts=50, end="natural", B=0, S=0):
    """CEEMDAN as emd_ref.py:220-259 defines it, on the C oracle.

    W: [E, N] standard normals (the reference draws default_rng(seed)).
    Stage 0 decomposes every W[e] (only the IMF counts are used); each stage
    k takes the first IMF of resid + eps*std(resid)*W[e] for every e (the raw
    noise, as the reference does), averages them, and subtracts the mean from
    resid; stops when resid is ~0 or every noise ran out of IMFs.
    Returns (imfs [K, N], resid, dict(noise_nimf, stage_sifts [K, E])).
    """
    x = _f64(x)
    W = np.ascontiguousarray(W, dtype=np.float64)
    E, N = W.shape
    noise_n = [emd_1d(W[e], tau=tau, max_sifts=max_sifts, max_imf=-1, min_extrema=4, end=end,
                      B=B, S=S)[0].shape[0] for e in range(E)]
    resid = x.copy()
    c_imfs, stage_sifts = [], []
    while True:
        if np.max(np.abs(resid)) < 1e-12 or np.std(resid) < 1e-10:
            break
        stage = len(c_imfs)
        beta = eps * np.std(resid)
        c = np.zeros(N)
        ns = []
        for e in range(E):
            imfs, _, st = emd_1d(resid + beta * W[e], tau=tau, max_sifts=max_sifts, max_imf=1,
                                 min_extrema=4, end=end, B=B, S=S)
            c += imfs[0]
            ns.append(st["per_mode_sifts"][0])
        c /= E
        c_imfs.append(c)
        stage_sifts.append(ns)
        resid = resid - c
        if all(stage >= n for n in noise_n):
            break
    return np.array(c_imfs), resid, dict(noise_nimf=np.array(noise_n, np.int32),
                                         stage_sifts=np.array(stage_sifts, np.int32))
