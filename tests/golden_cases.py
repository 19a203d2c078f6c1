"""Loader of the golden fixtures in tests/golden (written by tests/golden/make_fixtures.py)."""
from __future__ import annotations

import os

import numpy as np

from artefact_amd import Coefficient

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def loop_cases() -> list[dict]:
    """Each case: name, coefs, weight, pweight, iterations, max_w, max_h, mode, planes, objective."""
    with np.load(os.path.join(GOLDEN, "loop_cases.npz"), allow_pickle=False) as z:
        cases = []
        for i in range(int(z["count"])):
            p = f"c{i}_"
            C, W, H, iters, tv = (int(v) for v in z[p + "geometry"])
            params = z[p + "params"]
            coefs, planes = [], []
            for c in range(C):
                rw, rh, bw, bh, hs, vs = (int(v) for v in z[p + f"ch{c}_dims"])
                coefs.append(Coefficient(rw, rh, bw, bh, hs, vs, z[p + f"ch{c}_coefs"], z[p + f"ch{c}_quant"]))
                planes.append(z[p + f"ch{c}_out"])
            cases.append({
                "name": str(z[p + "name"]), "coefs": coefs, "weight": float(params[0]),
                "pweight": [float(v) for v in params[1:]], "iterations": iters, "max_w": W, "max_h": H,
                "mode": "tv" if tv else "tgv", "planes": planes, "objective": z[p + "objective"],
            })
        return cases


def dct_blocks() -> dict:
    with np.load(os.path.join(GOLDEN, "dct_blocks.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# The largest grid the decoder hands over: 16384 x 16384 4:2:0 (zune-jpeg/src/headers.rs:216-230).
# Its three planes take 3 GiB, past the 2^31 bytes up to which one 32-bit buffer offset
# addresses them all.
MAX_GRID = {"W": 16384, "H": 16384, "seed": 20240611, "quality": 75, "weight": 0.3, "pweight": [0.001] * 3,
            "iterations": 2}


def random_420(W: int, H: int, seed: int, quality: int) -> list:
    """Seeded 4:2:0 coefficients: DC in [-60, 60], the first nine AC terms in [-3, 3], the rest 0."""
    from artefact_amd import synth

    rng = np.random.default_rng(seed)
    out = []
    for c, (w, h, s) in enumerate([(W, H, 1), (W // 2, H // 2, 2), (W // 2, H // 2, 2)]):
        nb = (w // 8) * (h // 8)
        k = np.zeros((nb, 64), np.int16)
        k[:, 0] = rng.integers(-60, 61, nb, dtype=np.int16)
        k[:, 1:10] = rng.integers(-3, 4, (nb, 9), dtype=np.int16)
        out.append(Coefficient(w, h, w // 8, h // 8, s, s, k.reshape(-1), synth.quant_table(quality, c > 0)))
    return out


def max_grid_case() -> dict:
    """MAX_GRID with the SHA-256 of each output plane (f32 bytes) from the oracle."""
    import json

    with open(os.path.join(GOLDEN, "max_grid_420.json")) as f:
        fx = json.load(f)
    assert {k: fx[k] for k in MAX_GRID} == MAX_GRID, "fixture written for other parameters"
    return fx
