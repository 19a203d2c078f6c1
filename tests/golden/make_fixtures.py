"""Writes this directory's golden fixtures: small seeded workloads, their inputs, and the
oracle's output bits.  The oracle is the C restatement of pipeline_scalar, run with the norm
in f64.

    python tests/golden/make_fixtures.py

The reference's tests hold no golden vectors for this path, so these come from the oracle.
tests/test_oracle.py pins every oracle function against closed forms and brute force.  With
the fixtures fixed in bits, the oracle (tests/test_golden.py) and the HIP path
(tests/test_gpu_parity.py::test_golden_fixtures_bit_exact) are both checked against the same
stored bits.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from artefact_amd import synth  # noqa: E402

# (workload name, iterations, seed): every sampling path, TV and TGV, partial 64-pixel tiles.
LOOP_CASES = [
    ("tgv_420_40x24", 8, 101),
    ("tgv_444_24x16", 5, 102),
    ("tv_luma_24x16", 6, 103),
    ("tgv_422_32x16", 4, 104),
    ("tgv_440_16x32", 3, 105),
]
LOOP_FILE = os.path.join(HERE, "loop_cases.npz")
DCT_FILE = os.path.join(HERE, "dct_blocks.npz")


def loop_fixtures() -> dict:
    out = {}
    for i, (name, iters, seed) in enumerate(LOOP_CASES):
        wl = synth.workload(name, seed=seed)
        W, H = wl["max_w"], wl["max_h"]
        coefs = wl["coefs"]
        planes = oracle.compute(coefs, wl["weight"], wl["pweight"], iters, W, H, norm_mode="f64")
        total, terms = oracle.objective(coefs, wl["weight"], wl["pweight"], W, H, planes)
        p = f"c{i}_"
        out[p + "name"] = np.array(name)
        out[p + "params"] = np.array([wl["weight"], *wl["pweight"]], np.float32)
        out[p + "geometry"] = np.array([len(coefs), W, H, iters, 1 if wl["mode"] == "tv" else 0], np.int64)
        for c, k in enumerate(coefs):
            out[p + f"ch{c}_dims"] = np.array([k.rounded_px_w, k.rounded_px_h, k.block_w, k.block_h,
                                              k.horizontal_samp_factor, k.vertical_samp_factor], np.int64)
            out[p + f"ch{c}_coefs"] = k.dct_coefs
            out[p + f"ch{c}_quant"] = k.quant_table
            out[p + f"ch{c}_out"] = planes[c]
        out[p + "objective"] = np.array([total, *terms], np.float64)
    out["count"] = np.array(len(LOOP_CASES), np.int64)
    return out


def dct_fixtures() -> dict:
    rng = np.random.default_rng(7)
    blocks = np.concatenate([
        rng.normal(0.0, 60.0, (6, 64)),                          # image-like values
        rng.integers(-1024, 1024, (2, 64)).astype(np.float64),   # dequantised coefficients
        np.zeros((1, 64)), np.full((1, 64), 127.5),              # flat blocks
    ]).astype(np.float32)
    return {"blocks": blocks, "dct": oracle.dct8x8(blocks), "idct": oracle.idct8x8(blocks)}


def main() -> int:
    np.savez_compressed(LOOP_FILE, **loop_fixtures())
    np.savez_compressed(DCT_FILE, **dct_fixtures())
    print(LOOP_FILE, os.path.getsize(LOOP_FILE), "bytes")
    print(DCT_FILE, os.path.getsize(DCT_FILE), "bytes")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
