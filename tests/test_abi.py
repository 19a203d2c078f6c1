"""The C-ABI library loads and exports every entry point include/*.h declares
(CPU only: no compute call is made)."""
import ctypes
import os
import re

import numpy as np
import pytest

import artefact_amd
from artefact_amd import _lib
from artefact_amd.types import Coefficient

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(artefact_\w+)\s*\(", text))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert "artefact_hip_compute" in names
    assert "artefact_hip_plan_iterate" in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_binding_covers_every_declared_symbol():
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert declared_functions() <= bound


def test_strerror_without_gpu():
    lib = _lib.load()
    assert lib.artefact_hip_strerror(-3) == b"zero entry in a quantisation table"


def _coef(q0=16):
    q = np.full(64, 10, np.uint16)
    q[0] = q0
    return Coefficient(16, 16, 2, 2, 1, 1, np.zeros(256, np.int16), q)


def test_invalid_arguments_rejected_before_any_device_work():
    """Geometry and table errors are caught on the host (the reference would panic or
    divide by zero, compute_step_prob.rs:29)."""
    with pytest.raises(artefact_amd.ArtefactError) as e:
        artefact_amd.compute(1, [_coef(q0=0)], 0.3, [0.001], 5, 16, 16, 256)
    assert e.value.code == -3
    bad = _coef()
    bad.horizontal_samp_factor = 3
    with pytest.raises(artefact_amd.ArtefactError) as e:
        artefact_amd.compute(1, [bad], 0.3, [0.001], 5, 16, 16, 256)
    assert e.value.code == -2
    with pytest.raises(artefact_amd.ArtefactError) as e:
        artefact_amd.compute(1, [_coef()], 0.3, [0.001], 5, 8, 16, 128)  # component wider than the grid
    assert e.value.code == -2
    with pytest.raises(ValueError):
        artefact_amd.compute(1, [_coef()], 0.3, [0.001], 5, 16, 16, 100)


def test_coefficient_conversion_is_exact_or_refused():
    c = Coefficient(8, 8, 1, 1, 1, 1, np.arange(64, dtype=np.float32) - 32, np.ones(64, np.float32) * 3)
    assert c.dct_coefs.dtype == np.int16 and c.quant_table.dtype == np.uint16
    with pytest.raises(ValueError):
        Coefficient(8, 8, 1, 1, 1, 1, np.full(64, 0.5, np.float32), np.ones(64))


def test_missing_library_fails_loudly(monkeypatch):
    """No CPU fallback: without the HIP library every entry point raises instead of computing
    something else."""
    from artefact_amd import synth

    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libartefact_hip.so")
    monkeypatch.setattr(_lib, "_lib", None)
    wl = synth.workload("tgv_420_16x16", seed=1)
    W, H = wl["max_w"], wl["max_h"]
    with pytest.raises(RuntimeError, match="missing"):
        artefact_amd.compute(3, wl["coefs"], 0.3, [0.001] * 3, 2, W, H, W * H)
    with pytest.raises(RuntimeError, match="missing"):
        artefact_amd.Plan(wl["coefs"], 0.3, [0.001] * 3, 2, W, H)


def test_every_environment_switch_is_documented():
    """Each variable the library reads appears in INTEGRATION.md's table of switches."""
    import glob
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    names = set()
    for path in glob.glob(os.path.join(root, "artefact_amd", "csrc", "*")):
        with open(path, errors="replace") as f:
            names |= set(re.findall(r'getenv\("([A-Z_]+)"\)', f.read()))
    with open(os.path.join(root, "INTEGRATION.md")) as f:
        table = f.read().split("## 4. Environment switches")[1].split("\n## ")[0]
    assert names, "no getenv calls found"
    missing = sorted(n for n in names if f"`{n}" not in table)
    assert not missing, missing


def test_noalloc_rejects_grids_past_the_scratch():
    """artefact_hip_compute_noalloc checks the grid against its static scratch array before any
    device work: 4:2:0 2048x1536 needs about 188 MB, past the 64 MiB array."""
    from artefact_amd import synth
    from artefact_amd.compute import _channels, _params

    lib = _lib.load()
    wl = synth.workload("tgv_420_2048", seed=0)
    chans, keep = _channels(wl["coefs"])
    fake = ctypes.c_void_p(256)  # 16-byte aligned and never dereferenced: the size check comes first
    for c in range(3):
        chans[c].coefs = ctypes.cast(fake, ctypes.POINTER(ctypes.c_int16))
    outs = (ctypes.POINTER(ctypes.c_float) * 3)(*[ctypes.cast(fake, ctypes.POINTER(ctypes.c_float))
                                                  for _ in range(3)])
    p = _params(3, 0.3, [0.001] * 3, 100, wl["max_w"], wl["max_h"], "tgv", None, keep)
    rc = lib.artefact_hip_compute_noalloc(ctypes.byref(p), chans, ctypes.cast(outs, _lib.FLOAT_PP), None)
    assert rc == _lib.ARTEFACT_E_SCRATCH
    assert lib.artefact_hip_strerror(rc).startswith(b"grid too large")
