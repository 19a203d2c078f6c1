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
