"""Host-side checks that need no GPU: the C ABI library loads and exports every
entry point include/*.h declares, and the Python surface validates its inputs
the way the reference does (emd_par.py:281-299, eemd_stage.py:66-69)."""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"^\s*(?:int|void|const char\*)\s+(eemd_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    d = _declared()
    for n in ("eemd_ctx_create", "eemd_run_batch", "eemd_run_batch_dev", "eemd_noise"):
        assert n in d


def test_library_exports_every_declared_symbol():
    from eemd4cuda_amd.build import lib_path
    p = lib_path()
    if not p.exists():
        from eemd4cuda_amd.build import build
        build()
    lib = ctypes.CDLL(str(p))
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_params_default_without_gpu():
    from eemd4cuda_amd import _lib
    p = _lib.default_params()
    assert (p.fp64, p.end_mode, p.max_imf, p.max_sifts, p.min_extrema, p.block, p.halo) == \
        (1, 1, 10, 10, 4, 1024, 30)


def _c_layout(tmp_path, fields):
    """sizeof(eemd_params) and offsetof of each field, from the header by gcc."""
    src = tmp_path / "layout.c"
    body = "".join('  printf("%%zu\\n", offsetof(eemd_params, %s));\n' % f for f in fields)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "eemd_mi355x.h"\n'
                   'int main(void) {\n  printf("%zu\\n", sizeof(eemd_params));\n' + body +
                   '  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                           text=True).stdout.split()]
    return vals[0], vals[1:]


def test_params_struct_layout_matches_header(tmp_path):
    from eemd4cuda_amd import _lib
    names = [f[0] for f in _lib.Params._fields_]
    size, offs = _c_layout(tmp_path, names)
    assert ctypes.sizeof(_lib.Params) == size
    assert [getattr(_lib.Params, n).offset for n in names] == offs


def test_integration_stub_params_match_header(tmp_path):
    """The eemd_params mirror in INTEGRATION.md's ctypes stub has the header's layout."""
    ns = {}
    exec(_integration_params_source(), ns)
    P = ns["_Params"]
    names = [f[0] for f in P._fields_]
    size, offs = _c_layout(tmp_path, names)
    assert ctypes.sizeof(P) == size
    assert [getattr(P, n).offset for n in names] == offs


def _integration_params_source():
    txt = (ROOT / "INTEGRATION.md").read_text()
    block = re.search(r"```python\n(import ctypes as C.*?)```", txt, flags=re.S).group(1)
    a = block.index("class _Params")
    b = block.index("\n\n", a)
    return "import ctypes as C\n" + block[a:b] + "\n"


def test_decompose_input_validation():
    from eemd4cuda_amd.emd_par import decompose
    with pytest.raises(ValueError):
        decompose(np.ones(5), method="gpu")
    with pytest.raises(ValueError):
        decompose(np.ones((4, 4)), method="gpu")
    with pytest.raises(ValueError):
        decompose(np.ones(64), method="bogus")
    with pytest.raises(NotImplementedError):
        decompose(np.ones(64), method="cpu")


def test_run_eemd_error_codes():
    from eemd4cuda_amd.eemd_stage import EEMDError, run_eemd
    with pytest.raises(EEMDError) as ei:
        run_eemd(np.ones(5), method="gpu", E=1)
    assert ei.value.code == "eemd.bad_input"
    with pytest.raises(EEMDError) as ei:
        run_eemd(np.ones(64), method="cpu", E=1)
    assert ei.value.code == "eemd.backend"


def test_gpu_eemd_full_shape_guard():
    from eemd4cuda_amd.emd_gpu import gpu_eemd_full
    with pytest.raises(ValueError):
        gpu_eemd_full(np.zeros(100), E=4)


def test_signals_shapes():
    from eemd4cuda_amd import signals
    assert signals.three_tone().shape == (4096,)
    xs = signals.uneven_batch(8)
    assert all(4096 <= len(x) <= 262144 for x in xs)
    a = signals.audio_mix(1 << 12)
    assert a.shape == (4096,) and np.isfinite(a).all()
    b = signals.chirp_am_batch(2, 1024, seed=3)
    assert b.shape == (2, 1024)


def test_hip_runtime_choice(monkeypatch):
    """The HIP runtime the library binds to is explicit: EEMD_HIP_RUNTIME
    selects it, anything else is rejected."""
    from eemd4cuda_amd import _lib
    monkeypatch.setenv("EEMD_HIP_RUNTIME", "system")
    assert _lib._runtime_choice() == "system"
    monkeypatch.setenv("EEMD_HIP_RUNTIME", "torch")
    assert _lib._runtime_choice() == "torch"
    monkeypatch.setenv("EEMD_HIP_RUNTIME", "bogus")
    with pytest.raises(ValueError):
        _lib._runtime_choice()
