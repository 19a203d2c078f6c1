"""Writes the oracle's results for the full-size workloads, which take the oracle minutes:

* ``config_results.json``
  - ``tgv_420_2048``: 2048x1536 4:2:0 TGV, IJG q75, 100 iterations, seed 0.  SHA-256 of the
    three output planes (f32 loop, f64 gradient norm), four sample values per plane and the
    f64 objective with its terms.
  - ``tgv_444_4096``: 4096x3072 4:4:4 TGV, IJG q30, 100 iterations, seed 0.  The same for
    the f32 loop, plus the objective of the FP64 loop (``oracle.compute_f64``).
  - ``batch64``: the 64-image batch (``synth.batch_workload(64, seed=0)``) at 3 iterations:
    per image the SHA-256 of its planes (f32 loop, f64 gradient norm).
  - ``warm_420_2048_q50``: 2048x1536 4:2:0 TGV, IJG q50, seed 0 (the learned warm start's
    evaluation image): the FP64 loop run for 300 iterations (its converged reference), with
    its objective.
* ``fp64_444_4096_sub16.npz``: the FP64 loop's planes of ``tgv_444_4096``, every 16th row
  and column (offset 7), as float32 (plane ``p<c>``, shape (192, 256)).
* ``fp64_warm_420_2048_sub8.npz``: the converged FP64 planes of ``warm_420_2048_q50``, every
  8th row and column (offset 3), as float32 (shape (192, 256)), for PSNR estimates.

    python tests/golden/make_config_fixtures.py [tgv_420_2048] [tgv_444_4096] [batch64] [warm_420_2048_q50]

With no argument every part is written; a part named on the command line is recomputed and
the others are kept from the existing file.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from artefact_amd import synth  # noqa: E402

OUT = os.path.join(HERE, "config_results.json")
SUB = os.path.join(HERE, "fp64_444_4096_sub16.npz")
SUB_STEP, SUB_OFF = 16, 7
BATCH_ITERATIONS = 3


def digest(planes):
    return [hashlib.sha256(np.asarray(p, np.float32).astype("<f4").tobytes()).hexdigest() for p in planes]


def samples(planes):
    return [[float(np.asarray(p).reshape(-1)[i]) for i in (0, 12345, p.size // 2, p.size - 1)] for p in planes]


def single(name, fp64=False):
    wl = synth.workload(name, seed=0)
    W, H, it = wl["max_w"], wl["max_h"], wl["iterations"]
    t0 = time.time()
    out = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], it, W, H, norm_mode="f64")
    f, terms = oracle.objective(wl["coefs"], wl["weight"], wl["pweight"], W, H, out)
    rec = {"seed": 0, "width": W, "height": H, "iterations": it, "quality": wl["quality"], "mode": wl["mode"],
           "sha256": digest(out), "samples": samples(out), "objective": f, "terms": [float(t) for t in terms],
           "oracle_seconds": round(time.time() - t0, 1)}
    print(name, "f32", rec["oracle_seconds"], "s", flush=True)
    if fp64:
        t0 = time.time()
        out64 = oracle.compute_f64(wl["coefs"], wl["weight"], wl["pweight"], it, W, H)
        f64, t64 = oracle.objective(wl["coefs"], wl["weight"], wl["pweight"], W, H, out64)
        rec["fp64_objective"] = f64
        rec["fp64_terms"] = [float(t) for t in t64]
        rec["fp64_seconds"] = round(time.time() - t0, 1)
        sub = {f"p{c}": p.reshape(H, W)[SUB_OFF::SUB_STEP, SUB_OFF::SUB_STEP].astype(np.float32)
               for c, p in enumerate(out64)}
        np.savez_compressed(SUB, step=SUB_STEP, offset=SUB_OFF, **sub)
        rec["fp64_subsample"] = {"file": os.path.basename(SUB), "step": SUB_STEP, "offset": SUB_OFF}
        print(name, "fp64", rec["fp64_seconds"], "s", flush=True)
    return rec


def batch64():
    t0 = time.time()
    imgs = synth.batch_workload(64, seed=0, iterations=BATCH_ITERATIONS)
    rec = {"count": 64, "seed": 0, "iterations": BATCH_ITERATIONS, "images": []}
    for wl in imgs:
        out = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], BATCH_ITERATIONS, wl["max_w"], wl["max_h"],
                             norm_mode="f64")
        rec["images"].append({"name": wl["name"], "sha256": digest(out)})
    rec["oracle_seconds"] = round(time.time() - t0, 1)
    print("batch64", rec["oracle_seconds"], "s", flush=True)
    return rec


WARM_SUB = os.path.join(HERE, "fp64_warm_420_2048_sub8.npz")


def warm_converged():
    """The warm start's evaluation image: FP64 loop, 300 iterations (its converged reference)."""
    W, H, it = 2048, 1536, 300
    coefs = synth.encode(synth.synth_rgb(H, W, 0), subsampling="420", quality=50)
    t0 = time.time()
    out = oracle.compute_f64(coefs, 0.3, [0.001] * 3, it, W, H)
    f, terms = oracle.objective(coefs, 0.3, [0.001] * 3, W, H, out)
    sub = {f"p{c}": p.reshape(H, W)[3::8, 3::8].astype(np.float32) for c, p in enumerate(out)}
    np.savez_compressed(WARM_SUB, step=8, offset=3, **sub)
    rec = {"seed": 0, "width": W, "height": H, "quality": 50, "iterations": it, "fp64_objective": f,
           "fp64_terms": [float(t) for t in terms], "fp64_seconds": round(time.time() - t0, 1),
           "fp64_subsample": {"file": os.path.basename(WARM_SUB), "step": 8, "offset": 3}}
    print("warm_420_2048_q50", rec["fp64_seconds"], "s", flush=True)
    return rec


def main(argv):
    parts = argv or ["tgv_420_2048", "tgv_444_4096", "batch64", "warm_420_2048_q50"]
    res = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            res = json.load(f)
    for p in parts:
        if p == "tgv_420_2048":
            res[p] = single(p)
        elif p == "tgv_444_4096":
            res[p] = single(p, fp64=True)
        elif p == "batch64":
            res[p] = batch64()
        elif p == "warm_420_2048_q50":
            res[p] = warm_converged()
        else:
            raise SystemExit(f"unknown part {p}")
        with open(OUT, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
