"""artefact_hip_compute_noalloc: the loop on device buffers with every other buffer in the
library's static scratch array, against artefact_hip_compute (the same kernels, so the planes
must match bit for bit).  Two calls in a row, the second on a smaller grid, so the second runs
on what the first left in the scratch array.  Device buffers for the test's coefficients and
output planes come from hipMalloc through ctypes (the entry point itself allocates nothing)."""
import ctypes
import os

import numpy as np
import pytest

import artefact_amd
from artefact_amd import _lib, synth
from artefact_amd.compute import _channels, _params
from checks import bits_equal

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    lib = _lib.load()  # loads the HIP runtime the library links against
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = ctypes.CDLL(name)
            break
        except OSError:
            h = None
    assert h is not None, "HIP runtime not found"
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
        f.restype = ctypes.c_int
    return lib, h


class DeviceBuffer:
    def __init__(self, h, nbytes):
        self.h, self.ptr = h, ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes == self.nbytes
        assert self.h.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, dtype):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.h.hipFree(self.ptr)
            self.ptr = None


def noalloc(wl, iters):
    lib, h = hip()
    W, H = wl["max_w"], wl["max_h"]
    coefs = wl["coefs"]
    C = len(coefs)
    keep = []
    chans, k = _channels(coefs)
    keep += k
    dev = []
    try:
        for c, co in enumerate(coefs):
            b = DeviceBuffer(h, co.dct_coefs.size * 2)
            b.upload(co.dct_coefs.astype(np.int16))
            dev.append(b)
            chans[c].coefs = ctypes.cast(b.ptr, ctypes.POINTER(ctypes.c_int16))
        outs = [DeviceBuffer(h, W * H * 4) for _ in range(C)]
        dev += outs
        pp = (ctypes.POINTER(ctypes.c_float) * C)(*[ctypes.cast(o.ptr, ctypes.POINTER(ctypes.c_float)) for o in outs])
        p = _params(C, wl["weight"], wl["pweight"], iters, W, H, wl["mode"], None, keep)
        rc = lib.artefact_hip_compute_noalloc(ctypes.byref(p), chans, ctypes.cast(pp, _lib.FLOAT_PP), None)
        assert rc == 0, _lib.ERRORS.get(rc, rc)
        assert h.hipDeviceSynchronize() == 0
        return [o.download(np.float32) for o in outs]
    finally:
        for b in dev:
            b.free()


def test_noalloc_matches_compute_twice_in_a_row():
    cases = [("tgv_420_208x144", 11, 9), ("tv_luma_64x64", 12, 6)]  # the second grid is smaller
    for name, seed, iters in cases:
        wl = synth.workload(name, seed=seed)
        W, H = wl["max_w"], wl["max_h"]
        print(f"case noalloc {name} iters={iters}", flush=True)
        got = noalloc(wl, iters)
        want = artefact_amd.compute(len(wl["coefs"]), wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, W * H,
                                    mode=wl["mode"])
        for c, (a, b) in enumerate(zip(got, want)):
            assert bits_equal(a, b), f"{name} channel {c}: max abs diff {np.abs(a - b).max()}"

