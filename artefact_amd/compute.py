"""Host-side mirror of artefact-lib's reconstruction entry point.

``compute`` has the argument list and meaning of the crate-private
``compute`` of backend/artefact-lib/pipeline_scalar/mod.rs:17-26 and returns the
same thing (``nchannel`` f32 planes of ``max_rounded_px_count`` values, channel 0
still level-shifted).  It runs on the GPU through the C ABI of
include/artefact_hip.h; there is no CPU path.  ``Plan`` is the resident form
(coefficients stay in HBM; init, the iteration loop and the download are
separate calls) used for timing the loop alone.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .types import Coefficient


def _channels(coefs: Sequence[Coefficient]):
    arr = (_lib.artefact_channel * len(coefs))()
    keep = []
    for i, c in enumerate(coefs):
        d = np.ascontiguousarray(c.dct_coefs, dtype=np.int16)
        q = np.ascontiguousarray(c.quant_table, dtype=np.uint16)
        keep += [d, q]
        arr[i].rounded_w = c.rounded_px_w
        arr[i].rounded_h = c.rounded_px_h
        arr[i].block_w = c.block_w
        arr[i].block_h = c.block_h
        arr[i].h_samp = c.horizontal_samp_factor
        arr[i].v_samp = c.vertical_samp_factor
        arr[i].coefs = d.ctypes.data_as(ctypes.POINTER(ctypes.c_int16))
        arr[i].quant = q.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))
    return arr, keep


def _float_pp(planes):
    arr = (ctypes.POINTER(ctypes.c_float) * len(planes))()
    for i, p in enumerate(planes):
        arr[i] = p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    return arr


def _params(nchannel, weight, pweight, iterations, max_w, max_h, mode, init_override, keep):
    p = _lib.artefact_params()
    p.nchannel = nchannel
    p.max_w = max_w
    p.max_h = max_h
    p.weight = float(weight)
    pw = list(pweight) + [0.0] * (3 - len(pweight))
    for i in range(3):
        p.pweight[i] = float(pw[i])
    p.iterations = int(iterations)
    p.mode = _lib.ARTEFACT_MODE_TV if mode == "tv" else _lib.ARTEFACT_MODE_TGV
    if init_override is not None:
        planes = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in init_override]
        if len(planes) != nchannel or any(a.size != max_w * max_h for a in planes):
            raise ValueError("init_override needs nchannel planes of max_w*max_h values")
        keep += planes
        pp = _float_pp(planes)
        keep.append(pp)
        p.init_override = ctypes.cast(pp, _lib.FLOAT_PP)
    return p


def compute(
    nchannel: int,
    coefs: Sequence[Coefficient],
    weight: float,
    pweight: Sequence[float],
    iterations: int,
    max_rounded_px_w: int,
    max_rounded_px_h: int,
    max_rounded_px_count: int,
    *,
    mode: str = "tgv",
    init_override: Optional[Sequence[np.ndarray]] = None,
) -> list[np.ndarray]:
    """Runs the reconstruction loop on the GPU; see the module docstring.

    ``mode="tv"`` runs the TV-only loop (the second-order weight forced to 0).
    ``init_override`` replaces the decoded start image (one plane per channel).
    Raises ``ArtefactError`` on invalid geometry, a zero quantisation entry or a
    HIP failure (the reference panics in those cases).
    """
    if max_rounded_px_count != max_rounded_px_w * max_rounded_px_h:
        raise ValueError("max_rounded_px_count must equal max_rounded_px_w * max_rounded_px_h")
    if len(coefs) < nchannel:
        raise ValueError("fewer coefficient sets than channels")
    lib = _lib.load()
    keep: list = []
    chans, k2 = _channels(coefs[:nchannel])
    keep += k2
    p = _params(nchannel, weight, pweight, iterations, max_rounded_px_w, max_rounded_px_h, mode, init_override, keep)
    out = [np.empty(max_rounded_px_count, dtype=np.float32) for _ in range(nchannel)]
    outp = _float_pp(out)
    _lib.check(lib.artefact_hip_compute(ctypes.byref(p), chans, ctypes.cast(outp, _lib.FLOAT_PP)), "artefact_hip_compute")
    return out


class Plan:
    """Resident solve: coefficients uploaded once, then init / iterate / download."""

    def __init__(self, coefs: Sequence[Coefficient], weight: float, pweight: Sequence[float], iterations: int,
                 max_w: int, max_h: int, *, mode: str = "tgv", init_override=None, graph: bool = True):
        self._lib = _lib.load()
        self.nchannel = len(coefs)
        self.max_w, self.max_h = max_w, max_h
        self.M = [c.rounded_px_count for c in coefs]
        self._keep: list = []
        chans, k2 = _channels(coefs)
        self._keep += k2
        p = _params(self.nchannel, weight, pweight, iterations, max_w, max_h, mode, init_override, self._keep)
        h = ctypes.c_void_p()
        _lib.check(self._lib.artefact_hip_plan_create(ctypes.byref(p), chans, ctypes.byref(h)), "artefact_hip_plan_create")
        self._h = h
        _lib.check(self._lib.artefact_hip_plan_set_graph(self._h, 1 if graph else 0), "artefact_hip_plan_set_graph")

    def init(self) -> None:
        _lib.check(self._lib.artefact_hip_plan_init(self._h), "artefact_hip_plan_init")

    def iterate(self) -> None:
        _lib.check(self._lib.artefact_hip_plan_iterate(self._h), "artefact_hip_plan_iterate")

    def download(self) -> list[np.ndarray]:
        out = [np.empty(self.max_w * self.max_h, dtype=np.float32) for _ in range(self.nchannel)]
        _lib.check(self._lib.artefact_hip_plan_download(self._h, ctypes.cast(_float_pp(out), _lib.FLOAT_PP)),
                   "artefact_hip_plan_download")
        return out

    def download_cos(self) -> list[np.ndarray]:
        out = [np.empty(m, dtype=np.float32) for m in self.M]
        _lib.check(self._lib.artefact_hip_plan_download_cos(self._h, ctypes.cast(_float_pp(out), _lib.FLOAT_PP)),
                   "artefact_hip_plan_download_cos")
        return out

    def set_precision(self, precision: str) -> None:
        """"f32" (default, the reference's arithmetic bit for bit), "bf16x3" (MFMA projection
        transforms) or "f32_fast" (hardware square root and reciprocal in the gradient sweep,
        within a few ulp per term)."""
        code = {"f32": _lib.ARTEFACT_PRECISION_F32, "bf16x3": _lib.ARTEFACT_PRECISION_BF16X3,
                "f32_fast": _lib.ARTEFACT_PRECISION_F32_FAST}[precision]
        _lib.check(self._lib.artefact_hip_plan_set_precision(self._h, code), "artefact_hip_plan_set_precision")

    def tune(self) -> dict:
        """Autotunes strip height and projection kernel for this geometry; re-init after."""
        t = _lib.artefact_tuning()
        _lib.check(self._lib.artefact_hip_plan_tune(self._h, ctypes.byref(t)), "artefact_hip_plan_tune")
        return {"kg_rows": t.kg_rows, "project_kernel": {0: "per-channel", 1: "fused-lds", 2: "fused-lane"}[t.project_kernel],
                "ms_per_iteration": t.ms_per_iteration, "project_workgroups": t.project_workgroups,
                "xcd_bands": bool(t.xcd_bands)}

    def download_rgb(self, real_w: int, real_h: int) -> np.ndarray:
        """Output stage on the GPU (lib.rs:159-211): uint8 (real_h, real_w, 3)."""
        out = np.empty((real_h, real_w, 3), dtype=np.uint8)
        _lib.check(self._lib.artefact_hip_plan_download_rgb(self._h, real_w, real_h,
                                                            out.ctypes.data_as(_lib.U8_P)),
                   "artefact_hip_plan_download_rgb")
        return out

    def last_norms(self) -> np.ndarray:
        out = np.zeros(self.nchannel, dtype=np.float32)
        _lib.check(self._lib.artefact_hip_plan_last_norms(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   "artefact_hip_plan_last_norms")
        return out

    def first_gradient(self) -> list[np.ndarray]:
        out = [np.empty(self.max_w * self.max_h, dtype=np.float32) for _ in range(self.nchannel)]
        _lib.check(self._lib.artefact_hip_plan_first_gradient(self._h, ctypes.cast(_float_pp(out), _lib.FLOAT_PP)),
                   "artefact_hip_plan_first_gradient")
        return out

    def profile(self) -> dict:
        """Eager loop with per-kernel HIP events; consumes the initialised state."""
        out = np.zeros(3, dtype=np.float32)
        _lib.check(self._lib.artefact_hip_plan_profile(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   "artefact_hip_plan_profile")
        return {"k_gradient_ms": float(out[0]), "k_project_ms": float(out[1]), "loop_ms": float(out[2])}

    def set_timing(self, launches: int) -> None:
        """Record HIP events around the next ``launches`` iterate() launches (0: off)."""
        _lib.check(self._lib.artefact_hip_plan_set_timing(self._h, int(launches)), "artefact_hip_plan_set_timing")

    def timing(self) -> tuple[float, int]:
        """(summed duration in ms, number of timed launches) of the iterate() launches."""
        ms = ctypes.c_float(0.0)
        n = ctypes.c_uint32(0)
        _lib.check(self._lib.artefact_hip_plan_timing(self._h, ctypes.byref(ms), ctypes.byref(n)),
                   "artefact_hip_plan_timing")
        return float(ms.value), int(n.value)

    def objective(self) -> tuple[float, np.ndarray]:
        """f64 objective of the current iterate: (total, [TV, second-order, prob])."""
        t = np.zeros(3, dtype=np.float64)
        _lib.check(self._lib.artefact_hip_plan_objective(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                   "artefact_hip_plan_objective")
        return float(t.sum()), t

    @property
    def stream(self) -> int:
        return self._lib.artefact_hip_plan_stream(self._h) or 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.artefact_hip_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
