"""ctypes binding of libeemd_mi355x.so (the C ABI in include/eemd_mi355x.h).

There is no fallback: if the library or a GPU is missing, every entry point
raises.  One :class:`Engine` (device context) is created lazily per process.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import numpy as np

from .build import lib_path

EEMD_OK, EEMD_ERR_BAD_ARG, EEMD_ERR_DEVICE, EEMD_ERR_OOM = 0, 1, 2, 3
END_MODES = {"natural": 0, "mirror": 1, "learned": 2}
FLAG_MEMBER_IMFS = 1
FLAG_TRACE = 2
FLAG_PHASES = 4
FLAG_GRAPH = 8
FLAG_LEVEL_SUMS = 16
FLAG_TUNE_BLOCK = 32

class Params(C.Structure):
    _fields_ = [("fp64", C.c_int32), ("end_mode", C.c_int32), ("max_imf", C.c_int32),
                ("max_sifts", C.c_int32), ("min_extrema", C.c_int32), ("block", C.c_int32),
                ("halo", C.c_int32), ("threads", C.c_int32), ("members_per_task", C.c_int32),
                ("flags", C.c_int32), ("tau", C.c_double), ("seed", C.c_uint64),
                ("wave_windows", C.c_int32), ("member0", C.c_int32), ("signal0", C.c_int32)]


class EEMDBackendError(RuntimeError):
    """Device / allocation failure reported by the C ABI."""


_lib = None
_lock = threading.Lock()


def _runtime_choice() -> str:
    """Which HIP runtime the library binds to: EEMD_HIP_RUNTIME=torch | system.

    A process holds one HIP runtime.  PyTorch ships its own libamdhip64 with
    the same soname as ROCm's; whichever is loaded first serves the process,
    and a second copy loaded later starts a runtime that sees no GPU.
    'torch' (the default when PyTorch is installed) loads PyTorch's runtime
    first, so device memory from torch tensors and the engine's are one;
    'system' binds ROCm's runtime and leaves torch out of the process."""
    want = os.environ.get("EEMD_HIP_RUNTIME", "").strip().lower()
    if want not in ("", "torch", "system"):
        raise ValueError("EEMD_HIP_RUNTIME must be 'torch' or 'system'")
    if want == "system":
        return "system"
    try:
        import torch  # noqa: F401
    except ImportError:
        if want == "torch":
            raise RuntimeError("EEMD_HIP_RUNTIME=torch but PyTorch is not importable")
        return "system"
    return "torch"


def load(path: Path | None = None):
    """Load the shared library (raises if it has not been built).  The HIP
    runtime it binds to is chosen by _runtime_choice() before the load."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else lib_path()
    runtime = _runtime_choice()
    if not p.exists():
        raise RuntimeError(
            f"{p} not found: build it with `python -m eemd4cuda_amd.build` (hipcc, gfx950)")
    L = C.CDLL(str(p))
    vp, i32, i64, dbl, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_uint64
    L.eemd_params_default.argtypes = [C.POINTER(Params)]
    L.eemd_params_default.restype = None
    L.eemd_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.eemd_ctx_destroy.argtypes = [vp]
    L.eemd_last_error.argtypes = [vp]
    L.eemd_last_error.restype = C.c_char_p
    run_args = [vp, vp, vp, i32, i32, vp, vp, C.POINTER(Params), vp, vp, vp, vp, vp, vp, vp, vp]
    L.eemd_run_batch.argtypes = run_args
    L.eemd_run_batch_dev.argtypes = run_args + [vp]
    L.eemd_run_batch_dev_noalloc.argtypes = run_args + [vp]
    L.eemd_noise.argtypes = [vp, u64, i32, i32, i32, i64, vp]
    L.eemd_autotune.argtypes = [vp, vp, vp, i32, i32, vp, vp, C.POINTER(Params), i32,
                                C.POINTER(Params), C.POINTER(dbl)]
    L.eemd_last_timing.argtypes = [vp, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(i32),
                                   C.POINTER(i32), C.POINTER(i32)]
    L.eemd_last_phases.argtypes = [vp, C.POINTER(dbl)]
    L.eemd_graph_stats.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
    L.eemd_mem_info.argtypes = [vp, C.POINTER(u64), C.POINTER(u64)]
    for name in ("eemd_ctx_create", "eemd_ctx_destroy", "eemd_run_batch", "eemd_run_batch_dev",
                 "eemd_run_batch_dev_noalloc",
                 "eemd_noise", "eemd_last_timing", "eemd_last_phases", "eemd_autotune",
                 "eemd_graph_stats", "eemd_mem_info"):
        getattr(L, name).restype = C.c_int
    L._eemd_runtime = runtime
    _lib = L
    return L


def default_params() -> Params:
    p = Params()
    load().eemd_params_default(C.byref(p))
    return p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Engine:
    """A device context of the C ABI (device 0 by default)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.eemd_ctx_create(device, C.byref(h))
        if rc != EEMD_OK:
            raise EEMDBackendError(f"eemd_ctx_create(device={device}) failed with code {rc}: "
                                   "no usable HIP device")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.eemd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc == EEMD_OK:
            return
        msg = (self.lib.eemd_last_error(self.h) or b"").decode(errors="replace")
        if rc == EEMD_ERR_BAD_ARG:
            raise ValueError(f"{what}: {msg}")
        raise EEMDBackendError(f"{what} failed (code {rc}): {msg}")

    def run_batch(self, x: np.ndarray, offsets: np.ndarray, E: int, sigmas: np.ndarray,
                  prm: Params, noise: np.ndarray | None = None, member_imfs: bool = False,
                  trace: bool = False, per_mode: bool = True, graph: bool = False,
                  level_sums: bool = False):
        """Host-buffer EEMD of the concatenated signals (eemd_run_batch)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        sigmas = np.ascontiguousarray(sigmas, dtype=np.float64)
        nsig = len(offsets) - 1
        if len(sigmas) != nsig:
            raise ValueError("one sigma per signal is required")
        lens = np.diff(offsets)
        ntot = int(offsets[-1])
        if x.shape != (ntot,):
            raise ValueError("x must be the 1-D concatenation of the signals")
        if noise is not None:
            noise = np.ascontiguousarray(noise, dtype=np.float64)
            if noise.size != E * ntot:
                raise ValueError("noise must hold E normals per sample of every signal")
        mi = int(prm.max_imf)
        avg = np.zeros(mi * ntot)
        resid = np.zeros(ntot)
        n_modes = np.zeros(nsig, np.int32)
        nmode = np.zeros(nsig * E, np.int32)
        nsift = np.zeros(nsig * E, np.int32)
        pm = np.zeros(nsig * E * mi, np.int32) if per_mode else None
        flags = (FLAG_GRAPH if graph else 0) | (FLAG_LEVEL_SUMS if level_sums else 0)
        mim = None
        if member_imfs:
            flags |= FLAG_MEMBER_IMFS
            mim = np.zeros(E * mi * ntot, np.float64 if prm.fp64 else np.float32)
        tr = None
        if trace:
            flags |= FLAG_TRACE
            tr = np.zeros(nsig * E * mi * max(int(prm.max_sifts), 1) * 4, np.uint64)
        prm.flags = flags
        with _lock:
            rc = self.lib.eemd_run_batch(self.h, _ptr(x), _ptr(offsets), nsig, int(E),
                                         _ptr(sigmas), _ptr(noise), C.byref(prm), _ptr(avg),
                                         _ptr(resid), _ptr(n_modes), _ptr(nmode), _ptr(nsift),
                                         _ptr(pm), _ptr(mim), _ptr(tr))
        self._check(rc, "eemd_run_batch")
        out = dict(avg=avg, resid=resid, n_modes=n_modes, nmode=nmode.reshape(nsig, E),
                   nsift=nsift.reshape(nsig, E), lens=lens, offsets=offsets)
        if pm is not None:
            out["per_mode"] = pm.reshape(nsig, E, mi)
        if mim is not None:
            out["member_imfs"] = mim
        if tr is not None:
            out["trace"] = tr.reshape(nsig, E, mi, max(int(prm.max_sifts), 1), 4)
        out["timing"] = self.last_timing()
        return out

    def run_batch_dev(self, x_ptr: int, offsets: np.ndarray, E: int, sigmas: np.ndarray,
                      prm: Params, avg_ptr: int, resid_ptr: int, n_modes_ptr: int,
                      nmode_ptr: int, nsift_ptr: int, per_mode_ptr: int | None = None,
                      noise_ptr: int | None = None, stream: int | None = None,
                      noalloc: bool = False):
        """Device-buffer variant (eemd_run_batch_dev): pointers are device
        addresses.  noalloc=True calls eemd_run_batch_dev_noalloc, which takes
        its scratch from a static device array instead of allocating."""
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        sigmas = np.ascontiguousarray(sigmas, dtype=np.float64)
        nsig = len(offsets) - 1
        prm.flags = prm.flags & (FLAG_PHASES | FLAG_GRAPH | FLAG_LEVEL_SUMS | FLAG_TUNE_BLOCK)
        vp = C.c_void_p
        fn = self.lib.eemd_run_batch_dev_noalloc if noalloc else self.lib.eemd_run_batch_dev
        with _lock:
            rc = fn(
                self.h, vp(x_ptr), _ptr(offsets), nsig, int(E), _ptr(sigmas),
                vp(noise_ptr) if noise_ptr else None, C.byref(prm), vp(avg_ptr), vp(resid_ptr),
                vp(n_modes_ptr), vp(nmode_ptr), vp(nsift_ptr),
                vp(per_mode_ptr) if per_mode_ptr else None, None, None,
                vp(stream) if stream else None)
        self._check(rc, "eemd_run_batch_dev_noalloc" if noalloc else "eemd_run_batch_dev")

    def autotune(self, x_ptr: int, offsets: np.ndarray, E: int, sigmas: np.ndarray, prm: Params,
                 noise_ptr: int | None = None, reps: int = 2):
        """eemd_autotune on a device-resident batch: (best Params, kernel ms)."""
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        sigmas = np.ascontiguousarray(sigmas, dtype=np.float64)
        best = Params()
        ms = C.c_double()
        with _lock:
            rc = self.lib.eemd_autotune(self.h, C.c_void_p(x_ptr), _ptr(offsets), len(offsets) - 1,
                                        int(E), _ptr(sigmas),
                                        C.c_void_p(noise_ptr) if noise_ptr else None,
                                        C.byref(prm), int(reps), C.byref(best), C.byref(ms))
        self._check(rc, "eemd_autotune")
        return best, ms.value

    def noise(self, seed: int, signal: int, member0: int, E: int, n: int) -> np.ndarray:
        """The in-kernel Philox normals, materialised: [E, n]."""
        out = np.zeros((E, n))
        rc = self.lib.eemd_noise(self.h, C.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), int(signal),
                                 int(member0), int(E), int(n), _ptr(out))
        self._check(rc, "eemd_noise")
        return out

    def mem_info(self) -> tuple:
        """(free, total) device memory in bytes."""
        a, b = C.c_uint64(), C.c_uint64()
        self._check(self.lib.eemd_mem_info(self.h, C.byref(a), C.byref(b)), "eemd_mem_info")
        return a.value, b.value

    def graph_stats(self) -> tuple:
        """(graph captures, graph launches) of this context."""
        a, b = C.c_int64(), C.c_int64()
        self._check(self.lib.eemd_graph_stats(self.h, C.byref(a), C.byref(b)), "eemd_graph_stats")
        return a.value, b.value

    def last_phases(self) -> list:
        out = (C.c_double * 4)()
        rc = self.lib.eemd_last_phases(self.h, out)
        self._check(rc, "eemd_last_phases")
        return list(out)

    def last_timing(self) -> dict:
        a, b = C.c_double(), C.c_double()
        t, g, n = C.c_int32(), C.c_int32(), C.c_int32()
        self.lib.eemd_last_timing(self.h, C.byref(a), C.byref(b), C.byref(t), C.byref(g),
                                  C.byref(n))
        return dict(sift_ms=a.value, mean_ms=b.value, threads=t.value, grid=g.value,
                    ntasks=n.value)


_engine = None


def engine() -> Engine:
    """The process-wide device context (created on first use)."""
    global _engine
    with _lock:
        if _engine is None:
            _engine = Engine(0)
        return _engine
