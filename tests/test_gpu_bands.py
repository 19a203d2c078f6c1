"""Projection workgroups by XCD row band (ARTEFACT_XCD_BANDS).

The bands only change which workgroup projects which blocks, not the arithmetic, so the loop
must give the same bits with them on and off, and the f32 loop the oracle's bits.  The grids
are large enough for at least eight projection workgroups (with fewer, the lane kernel does
not band), and cover the lane kernel (4:2:0, 4:4:4, luma), the LDS kernel, the per-channel
kernel (4:2:2) and the BF16 MFMA kernel.
"""
import numpy as np
import pytest

import artefact_amd
import oracle
from artefact_amd import synth

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def run(wl, iters, precision):
    W, H = wl["max_w"], wl["max_h"]
    with artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, mode=wl["mode"]) as plan:
        plan.set_precision(precision)
        plan.init()
        plan.iterate()
        return plan.download()


@pytest.mark.parametrize("name,iters,precision,kernel", [
    ("tgv_420_208x144", 6, "f32", "lane"),
    ("tgv_420_208x144", 6, "f32", "lds"),
    ("tgv_444_160x96", 5, "f32", "lane"),
    ("tv_luma_256x256", 5, "f32", "lane"),
    ("tgv_422_256x128", 5, "f32", "lane"),
    ("tgv_444_160x96", 5, "bf16x3", "lane"),
])
def test_bands_give_the_same_bits(name, iters, precision, kernel, monkeypatch):
    monkeypatch.setenv("ARTEFACT_PROJECT_KERNEL", kernel)
    wl = synth.workload(name, seed=41)
    print(f"case xcd bands {name} {precision} {kernel} iters={iters}", flush=True)
    out = {}
    for bands in ("0", "1"):
        monkeypatch.setenv("ARTEFACT_XCD_BANDS", bands)
        out[bands] = run(wl, iters, precision)
    for a, b in zip(out["0"], out["1"]):
        assert bits_equal(a, b)
    if precision == "f32":
        ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], iters, wl["max_w"], wl["max_h"],
                             norm_mode="f64")
        for a, b in zip(out["1"], ref):
            assert bits_equal(a, b)
