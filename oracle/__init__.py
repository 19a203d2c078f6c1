"""CPU oracle for artefact-lib's reconstruction loop -- TEST INFRASTRUCTURE ONLY.

ctypes front end of ``liboracle.so`` (oracle/oracle.c, a restatement of
backend/artefact-lib/pipeline_scalar).  Only tests/, ``__graft_entry__.smoke()``
and bench.py's ``cpu_baseline`` leg use it, and only as the checker / the timed
CPU baseline.  The product package (artefact_amd) never imports it.

Parity status per function: see the header of oracle/oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

INIT_MODES = {"corrected": 0, "scalar_quirk": 1}
NORM_MODES = {"reference": 0, "f64": 1}


class oracle_channel(ctypes.Structure):
    _fields_ = [
        ("rounded_w", ctypes.c_uint32),
        ("rounded_h", ctypes.c_uint32),
        ("block_w", ctypes.c_uint32),
        ("block_h", ctypes.c_uint32),
        ("h_samp", ctypes.c_uint32),
        ("v_samp", ctypes.c_uint32),
        ("coefs", ctypes.POINTER(ctypes.c_float)),
        ("quant", ctypes.POINTER(ctypes.c_float)),
    ]


_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_FPP = ctypes.POINTER(_F)
_DPP = ctypes.POINTER(_D)
_CH = ctypes.POINTER(oracle_channel)
_u32 = ctypes.c_uint32

_SIGS = {
    "oracle_compute_f32": (ctypes.c_int, [_u32, _CH, ctypes.c_float, _F, _u32, _u32, _u32, ctypes.c_int,
                                          ctypes.c_int, _FPP, _FPP, _D, _F]),
    "oracle_compute_f64": (ctypes.c_int, [_u32, _CH, ctypes.c_float, _F, _u32, _u32, _u32, ctypes.c_int,
                                          _DPP, _DPP, _D]),
    "oracle_init_f32": (ctypes.c_int, [_CH, _u32, _u32, ctypes.c_int, _F]),
    "oracle_gradient_f32": (ctypes.c_int, [_u32, _CH, ctypes.c_float, _F, _u32, _u32, _FPP, _FPP, _FPP, _D, _D]),
    "oracle_projection_f32": (ctypes.c_int, [_CH, _u32, _u32, _F, _F]),
    "oracle_dct8x8_f32": (None, [_F, ctypes.c_uint64]),
    "oracle_idct8x8_f32": (None, [_F, ctypes.c_uint64]),
    "oracle_dct8x8_f64": (None, [_D, ctypes.c_uint64]),
    "oracle_idct8x8_f64": (None, [_D, ctypes.c_uint64]),
    "oracle_boxing_f32": (None, [_F, _F, _u32, _u32, _u32]),
    "oracle_unboxing_f32": (None, [_F, _F, _u32, _u32, _u32]),
    "oracle_midpoint_f32": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
    "oracle_fista_scalars": (None, [_u32, _u32, _F, _F, _F]),
    "oracle_objective": (ctypes.c_double, [_u32, _CH, ctypes.c_float, _F, _u32, _u32, _FPP, _D]),
    "oracle_objective_f64planes": (ctypes.c_double, [_u32, _CH, ctypes.c_float, _F, _u32, _u32, _DPP, _D]),
    "oracle_output_rgb": (None, [_u32, _FPP, _u32, _u32, _u32, ctypes.POINTER(ctypes.c_uint8)]),
    "oracle_set_threads": (None, [ctypes.c_int]),
}

_lib = None


def build() -> None:
    """Compiles liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _fptr(a):
    return a.ctypes.data_as(_F)


def _dptr(a):
    return a.ctypes.data_as(_D)


def _chan_array(coefs):
    arr = (oracle_channel * len(coefs))()
    keep = []
    for i, c in enumerate(coefs):
        d = np.ascontiguousarray(np.asarray(c.dct_coefs).reshape(-1), dtype=np.float32)
        q = np.ascontiguousarray(np.asarray(c.quant_table).reshape(-1), dtype=np.float32)
        keep += [d, q]
        arr[i].rounded_w = c.rounded_px_w
        arr[i].rounded_h = c.rounded_px_h
        arr[i].block_w = c.block_w
        arr[i].block_h = c.block_h
        arr[i].h_samp = c.horizontal_samp_factor
        arr[i].v_samp = c.vertical_samp_factor
        arr[i].coefs = _fptr(d)
        arr[i].quant = _fptr(q)
    return arr, keep


def _fpp(planes):
    arr = (_F * len(planes))()
    for i, p in enumerate(planes):
        arr[i] = _fptr(p)
    return arr


def _dpp(planes):
    arr = (_D * len(planes))()
    for i, p in enumerate(planes):
        arr[i] = _dptr(p)
    return arr


def _pw(pweight):
    p = list(pweight) + [0.0] * (3 - len(pweight))
    return np.asarray(p[:3], dtype=np.float32)


def _check(rc, where):
    if rc != 0:
        raise ValueError(f"{where} failed with code {rc}")


def compute(coefs, weight, pweight, iterations, max_w, max_h, *, nchannel=None, init_mode="corrected",
            norm_mode="reference", init_override=None, return_norms=False, return_time=False, threads=1):
    """f32 loop of pipeline_scalar/mod.rs:17-78.  Returns nchannel planes (N f32 each).

    ``threads`` > 1 runs the loops the reference runs over channels with rayon (the FISTA
    update, mod.rs:54-59) one thread per channel; compute_step stays serial, as there."""
    lib = load()
    lib.oracle_set_threads(int(threads))
    C = nchannel or len(coefs)
    chans, keep = _chan_array(coefs[:C])
    N = max_w * max_h
    out = [np.empty(N, dtype=np.float32) for _ in range(C)]
    pw = _pw(pweight)
    ov = None
    if init_override is not None:
        ovp = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in init_override]
        keep += ovp
        ov = _fpp(ovp)
    secs = ctypes.c_double(0.0)
    norms = np.zeros(max(iterations, 1) * C, dtype=np.float32)
    rc = lib.oracle_compute_f32(C, chans, float(weight), _fptr(pw), iterations, max_w, max_h, INIT_MODES[init_mode],
                                NORM_MODES[norm_mode], ov, _fpp(out), ctypes.byref(secs), _fptr(norms))
    _check(rc, "oracle_compute_f32")
    res = [out]
    if return_norms:
        res.append(norms[: iterations * C].reshape(iterations, C))
    if return_time:
        res.append(secs.value)
    return res[0] if len(res) == 1 else tuple(res)


def compute_f64(coefs, weight, pweight, iterations, max_w, max_h, *, nchannel=None, init_mode="corrected",
                init_override=None, return_time=False):
    """The same loop in f64 (norm in f64), the FP64 reference for the precision form."""
    lib = load()
    lib.oracle_set_threads(1)
    C = nchannel or len(coefs)
    chans, keep = _chan_array(coefs[:C])
    N = max_w * max_h
    out = [np.empty(N, dtype=np.float64) for _ in range(C)]
    ov = None
    if init_override is not None:
        ovp = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in init_override]
        keep += ovp
        ov = _dpp(ovp)
    secs = ctypes.c_double(0.0)
    pw = _pw(pweight)
    rc = lib.oracle_compute_f64(C, chans, float(weight), _fptr(pw), iterations, max_w, max_h, INIT_MODES[init_mode],
                                ov, _dpp(out), ctypes.byref(secs))
    _check(rc, "oracle_compute_f64")
    return (out, secs.value) if return_time else out


def init_plane(coef, max_w, max_h, init_mode="corrected"):
    lib = load()
    chans, keep = _chan_array([coef])
    out = np.empty(max_w * max_h, dtype=np.float32)
    _check(lib.oracle_init_f32(chans, max_w, max_h, INIT_MODES[init_mode], _fptr(out)), "oracle_init_f32")
    return out


def gradient(coefs, weight, pweight, max_w, max_h, fdata, cos):
    """prob + TV + second-order gradient (compute_step.rs:21-48) for given planes and cos."""
    lib = load()
    C = len(fdata)
    chans, keep = _chan_array(coefs[:C])
    fd = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in fdata]
    cs = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in cos]
    g = [np.empty(max_w * max_h, dtype=np.float32) for _ in range(C)]
    tv = ctypes.c_double()
    tv2 = ctypes.c_double()
    pw = _pw(pweight)
    rc = lib.oracle_gradient_f32(C, chans, float(weight), _fptr(pw), max_w, max_h, _fpp(fd), _fpp(cs), _fpp(g),
                                 ctypes.byref(tv), ctypes.byref(tv2))
    _check(rc, "oracle_gradient_f32")
    return g, tv.value, tv2.value


def projection(coef, max_w, max_h, fdata):
    """compute_projection.rs:8-122 on one channel; returns (projected plane, cos)."""
    lib = load()
    chans, keep = _chan_array([coef])
    fd = np.array(fdata, dtype=np.float32).reshape(-1).copy()
    cos = np.empty(coef.rounded_px_count, dtype=np.float32)
    _check(lib.oracle_projection_f32(chans, max_w, max_h, _fptr(fd), _fptr(cos)), "oracle_projection_f32")
    return fd, cos


def _blocks(fn, blocks, dtype):
    b = np.array(blocks, dtype=dtype).reshape(-1, 64).copy()
    fn(b.ctypes.data_as(_F if dtype == np.float32 else _D), b.shape[0])
    return b


def dct8x8(blocks, dtype=np.float32):
    lib = load()
    return _blocks(lib.oracle_dct8x8_f32 if dtype == np.float32 else lib.oracle_dct8x8_f64, blocks, dtype)


def idct8x8(blocks, dtype=np.float32):
    lib = load()
    return _blocks(lib.oracle_idct8x8_f32 if dtype == np.float32 else lib.oracle_idct8x8_f64, blocks, dtype)


def boxing(plane, rw, bw, bh):
    lib = load()
    src = np.ascontiguousarray(plane, dtype=np.float32).reshape(-1)
    out = np.empty(bw * bh * 64, dtype=np.float32)
    lib.oracle_boxing_f32(_fptr(src), _fptr(out), rw, bw, bh)
    return out


def unboxing(blocks, rw, bw, bh):
    lib = load()
    src = np.ascontiguousarray(blocks, dtype=np.float32).reshape(-1)
    out = np.empty(bw * bh * 64, dtype=np.float32)
    lib.oracle_unboxing_f32(_fptr(src), _fptr(out), rw, bw, bh)
    return out


def midpoint(a, b):
    return load().oracle_midpoint_f32(float(a), float(b))


def fista_scalars(iterations, pixel_count):
    lib = load()
    nt = np.empty(max(iterations, 1), dtype=np.float32)
    fa = np.empty(max(iterations, 1), dtype=np.float32)
    st = ctypes.c_float()
    lib.oracle_fista_scalars(iterations, pixel_count, _fptr(nt), _fptr(fa), ctypes.byref(st))
    return nt[:iterations], fa[:iterations], st.value


def objective(coefs, weight, pweight, max_w, max_h, planes):
    """Build-defined FP64 objective (oracle/oracle.c objective_impl); returns (total, [tv, tgv, prob])."""
    lib = load()
    C = len(planes)
    chans, keep = _chan_array(coefs[:C])
    terms = np.zeros(3, dtype=np.float64)
    pw = _pw(pweight)
    if np.asarray(planes[0]).dtype == np.float64:
        ps = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in planes]
        tot = lib.oracle_objective_f64planes(C, chans, float(weight), _fptr(pw), max_w, max_h, _dpp(ps), _dptr(terms))
    else:
        ps = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in planes]
        tot = lib.oracle_objective(C, chans, float(weight), _fptr(pw), max_w, max_h, _fpp(ps), _dptr(terms))
    return tot, terms


def output_rgb(planes, max_w, real_w, real_h):
    """Output stage of backend/artefact-lib/lib.rs:159-211 on loop planes (luma level-shifted).

    Returns uint8 (real_h, real_w, 3); grey (R = G = B) unless there are three planes.
    """
    lib = load()
    ps = [np.ascontiguousarray(a, dtype=np.float32).reshape(-1) for a in planes]
    out = np.empty((real_h, real_w, 3), dtype=np.uint8)
    lib.oracle_output_rgb(len(ps), _fpp(ps), max_w, real_w, real_h,
                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out


def process(coefs, real_w, real_h, weight, pweight, iterations, separate_components=False, *,
            norm_mode="reference", init_mode="corrected"):
    """Artefact::process after decoding (backend/artefact-lib/lib.rs:105-211).

    ``weight``, ``pweight`` and ``iterations`` are 3-element lists (ValueCollection::to_slice,
    lib.rs:41-46).  Three components and not ``separate_components``: one joint loop with
    weight[0] and iterations[0] (:121-132).  Otherwise one single-channel loop per component with
    weight[c] and iterations[c]; each receives the whole pweight array and so uses pweight[0]
    (:134-149 with compute_step.rs:19-30).  Returns uint8 (real_h, real_w, 3).
    """
    W = max(c.rounded_px_w for c in coefs)
    H = max(c.rounded_px_h for c in coefs)
    if len(coefs) == 3 and not separate_components:
        planes = compute(coefs, weight[0], pweight, iterations[0], W, H, norm_mode=norm_mode, init_mode=init_mode)
    else:
        planes = [compute([c], weight[i], pweight, iterations[i], W, H, norm_mode=norm_mode,
                          init_mode=init_mode)[0] for i, c in enumerate(coefs)]
    return output_rgb(planes, W, real_w, real_h)
