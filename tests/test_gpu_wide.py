"""The gradient sweep with 64-bit addressing (per-row buffer resources).

Grids whose C planes reach 2^31 bytes run k_gradient's WIDE instance, whose offsets cannot
wrap.  ARTEFACT_WIDE_ADDRESSING=1 selects it for any plan, so the small parity cases check
that it is bit-identical to the oracle.  The largest grid the decoder accepts (16384 x 16384
4:2:0, 3 GiB of planes) is checked against the oracle's output hashes
(tests/golden/make_max_grid.py).
"""
import hashlib

import numpy as np
import pytest

import artefact_amd
import oracle
from artefact_amd import synth
from golden_cases import MAX_GRID, max_grid_case, random_420

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def run_gpu(wl, iters):
    W, H = wl["max_w"], wl["max_h"]
    return artefact_amd.compute(len(wl["coefs"]), wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, W * H,
                                mode=wl["mode"])


@pytest.mark.parametrize("name,iters", [("tgv_420_208x144", 9), ("tgv_444_72x40", 7), ("tgv_422_80x40", 6),
                                        ("tgv_440_48x64", 6), ("tv_luma_64x64", 8), ("tgv_420_16x16", 5)])
def test_wide_addressing_bit_exact(name, iters, monkeypatch):
    monkeypatch.setenv("ARTEFACT_WIDE_ADDRESSING", "1")
    wl = synth.workload(name, seed=31)
    print(f"case wide {name} {wl['max_w']}x{wl['max_h']} iters={iters}", flush=True)
    out = run_gpu(wl, iters)
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], iters, wl["max_w"], wl["max_h"],
                         norm_mode="f64")
    for a, b in zip(out, ref):
        assert bits_equal(a, b)


def test_wide_fast_mode_same_bits(monkeypatch):
    """The fast mode changes the arithmetic, not the addressing: WIDE gives the same bits."""
    wl = synth.workload("tgv_420_208x144", seed=32)
    print("case wide fast tgv_420_208x144", flush=True)
    monkeypatch.setenv("ARTEFACT_PRECISION", "f32_fast")
    plain = run_gpu(wl, 8)
    monkeypatch.setenv("ARTEFACT_WIDE_ADDRESSING", "1")
    wide = run_gpu(wl, 8)
    for a, b in zip(plain, wide):
        assert bits_equal(a, b)


def test_largest_grid_matches_oracle_hashes():
    m = MAX_GRID
    fx = max_grid_case()
    print(f"case max grid {m['W']}x{m['H']} 4:2:0 iters={m['iterations']}", flush=True)
    coefs = random_420(m["W"], m["H"], m["seed"], m["quality"])
    out = artefact_amd.compute(3, coefs, m["weight"], m["pweight"], m["iterations"], m["W"], m["H"], m["W"] * m["H"])
    for c, p in enumerate(out):
        flat = np.asarray(p, np.float32).reshape(-1)
        samples = [float(flat[i]) for i in (0, 12345, flat.size // 2, flat.size - 1)]
        assert samples == fx["samples"][c], (c, samples)
        assert hashlib.sha256(flat.astype("<f4").tobytes()).hexdigest() == fx["sha256"][c], c
