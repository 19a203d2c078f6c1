"""Writes tests/golden/max_grid_420.json: the oracle's output hashes for the largest grid.

The case (tests/golden_cases.py MAX_GRID) is 16384 x 16384 4:2:0 with seeded random
coefficients, 2 TGV iterations, f64 gradient norms.  The oracle needs about 2 minutes and
25 GB for it, so the GPU test compares against these hashes instead of running it.

    python tests/golden/make_max_grid.py
"""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from golden_cases import MAX_GRID, random_420  # noqa: E402


def main() -> None:
    m = MAX_GRID
    coefs = random_420(m["W"], m["H"], m["seed"], m["quality"])
    t0 = time.time()
    planes = oracle.compute(coefs, m["weight"], m["pweight"], m["iterations"], m["W"], m["H"], norm_mode="f64")
    dt = time.time() - t0
    fx = dict(m)
    fx["sha256"] = [hashlib.sha256(p.astype("<f4").tobytes()).hexdigest() for p in planes]
    fx["samples"] = [[float(p.reshape(-1)[i]) for i in (0, 12345, p.size // 2, p.size - 1)] for p in planes]
    fx["oracle_seconds"] = round(dt, 1)
    with open(os.path.join(HERE, "max_grid_420.json"), "w") as f:
        json.dump(fx, f, indent=1)
    print(json.dumps(fx, indent=1))


if __name__ == "__main__":
    main()
