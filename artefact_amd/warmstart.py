"""Learned warm start: a residual CNN whose prediction starts the reconstruction loop.

The network (csrc/cnn.hip) maps a decoded plane and its quantisation-step map (the channel's
64-entry table tiled over its 8x8 blocks) to a prediction of the converged plane, at the
component's own resolution.  Training pairs come from the solver itself: the decoded start
image of a synthetic crop (already inside the quantisation box) and the plain TGV loop's output
after ``target_iterations`` iterations (for subsampled chroma, its 2x2 means: what the
projection keeps of a prediction replicated onto the grid).  Forward, backward and Adam run on
the GPU.  ``warm_solve`` starts a plan from the projected prediction, runs its iterations and
falls back to the plain loop when the warm result's objective is higher than the plain
result's (the objective decides).  No reference counterpart.
"""
from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass

import numpy as np

from . import _lib, synth
from .batch import Batch, BatchImage
from .compute import Plan

_FP = ctypes.POINTER(ctypes.c_float)
WIDTHS = (2, 32, 32, 32, 32, 1)
WEIGHTS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_weights.npz")
# What the weights file's network was trained on (stored with the weights, checked on load).
INPUT_FEATURES = "plane/64 and tiled quantisation table/64, components at their own resolution"


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised flat parameters (per layer w[cout][cin][3][3], then b[cout]); the last
    layer is scaled by 0.1 so that the untrained network is close to the identity map."""
    rng = np.random.default_rng(seed)
    out = []
    for li, (cin, cout) in enumerate(zip(WIDTHS[:-1], WIDTHS[1:])):
        w = rng.normal(0.0, np.sqrt(2.0 / (cin * 9)), size=(cout, cin, 3, 3))
        if li == len(WIDTHS) - 2:
            w *= 0.1
        out += [w.reshape(-1), np.zeros(cout)]
    return np.concatenate(out).astype(np.float32)


def split_params(flat: np.ndarray):
    """Flat vector -> ([w_l (cout, cin, 3, 3)], [b_l]) in float64."""
    ws, bs, o = [], [], 0
    for cin, cout in zip(WIDTHS[:-1], WIDTHS[1:]):
        n = cout * cin * 9
        ws.append(flat[o:o + n].astype(np.float64).reshape(cout, cin, 3, 3))
        o += n
        bs.append(flat[o:o + cout].astype(np.float64))
        o += cout
    return ws, bs


def _fp(a):
    return a.ctypes.data_as(_FP)


class WarmStartNet:
    def __init__(self, params: np.ndarray | None = None):
        self._lib = _lib.load()
        n = self._lib.artefact_cnn_param_count()
        p = init_params() if params is None else np.ascontiguousarray(params, dtype=np.float32)
        if p.size != n:
            raise ValueError(f"expected {n} parameters, got {p.size}")
        h = ctypes.c_void_p()
        _lib.check(self._lib.artefact_cnn_create(_fp(p), ctypes.byref(h)), "artefact_cnn_create")
        self._h = h
        self.n = n

    @classmethod
    def load(cls, path: str = WEIGHTS_FILE) -> "WarmStartNet":
        with np.load(path, allow_pickle=False) as f:
            if str(f.get("input_features", "")) != INPUT_FEATURES:
                raise ValueError(f"{path}: weights trained on other input features")
            return cls(f["params"])

    def save(self, path: str, **meta) -> None:
        np.savez(path, params=self.params(), **{k: np.asarray(v) for k, v in meta.items()})

    def params(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float32)
        _lib.check(self._lib.artefact_cnn_get_params(self._h, _fp(out)), "artefact_cnn_get_params")
        return out

    def forward(self, x0: np.ndarray, quant) -> np.ndarray:
        """Predictions (image units) for planes x0 (B, H, W) with quantisation tables (B, 64)."""
        x = np.ascontiguousarray(x0, dtype=np.float32)
        q = _tables(quant, x.shape[0])
        out = np.empty_like(x)
        B, H, W = x.shape
        _lib.check(self._lib.artefact_cnn_forward(self._h, _fp(x), _fp(q), B, H, W, _fp(out)), "artefact_cnn_forward")
        return out

    def gradients(self, x0, quant, target):
        """(loss, flat parameter gradients) of the mean squared error in units of 1/64."""
        x = np.ascontiguousarray(x0, dtype=np.float32)
        q = _tables(quant, x.shape[0])
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.empty(self.n, dtype=np.float32)
        loss = ctypes.c_double()
        B, H, W = x.shape
        _lib.check(self._lib.artefact_cnn_gradients(self._h, _fp(x), _fp(q), _fp(t), B, H, W, ctypes.byref(loss),
                                                    _fp(g)), "artefact_cnn_gradients")
        return loss.value, g

    def set_dataset(self, x0, quant, target) -> None:
        x = np.ascontiguousarray(x0, dtype=np.float32)
        q = _tables(quant, x.shape[0])
        t = np.ascontiguousarray(target, dtype=np.float32)
        B, H, W = x.shape
        _lib.check(self._lib.artefact_cnn_set_dataset(self._h, _fp(x), _fp(q), _fp(t), B, H, W),
                   "artefact_cnn_set_dataset")
        self.count = B

    def train_step(self, indices, lr: float) -> float:
        idx = np.ascontiguousarray(indices, dtype=np.uint32)
        loss = ctypes.c_double()
        _lib.check(self._lib.artefact_cnn_train_step(self._h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                     idx.size, float(lr), ctypes.byref(loss)),
                   "artefact_cnn_train_step")
        return loss.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.artefact_cnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _tables(quant, n):
    q = np.ascontiguousarray(quant, dtype=np.float32).reshape(-1, 64)
    if q.shape[0] != n:
        raise ValueError(f"expected {n} quantisation tables of 64 entries, got {q.shape[0]}")
    return q


def quant_table(coef) -> np.ndarray:
    """The component's quantisation table (natural order) as the network's feature input."""
    return np.asarray(coef.quant_table, dtype=np.float32).reshape(64)


def _encode_crop(job):
    """One synthetic crop, encoded (runs in a worker process; no GPU use)."""
    size, img_seed, quality = job
    return synth.encode(synth.synth_rgb(size, size, img_seed), subsampling="420", quality=quality)


SAMPLES_PER_CROP = 6  # four luma quarters and the two chroma components of a 2 size x 2 size crop


def training_pairs(count: int, size: int = 256, seed: int = 0, target_iterations: int = 300,
                   qualities=(20, 90), weight: float = 0.3, pweight: float = 0.001, log=print,
                   workers: int = 8):
    """``count`` synthetic 4:2:0 crops of 2 size x 2 size at qualities drawn uniformly in the
    range, each giving six samples of size x size at the resolution the network sees in
    inference: the four quarters of the luma plane, and each chroma component at its own
    (half) resolution.  Returns (x0, quant, target): x0 the decoded start plane (a chroma
    component before its upsampling), quant its quantisation table (per sample, 64 entries),
    target the plain TGV loop's output (for chroma, the 2x2 means of the solved grid plane), all
    float32.  The crops are synthesised and encoded by ``workers`` spawned processes (host work
    only); the start images and targets are solved on the GPU by this process."""
    import multiprocessing

    rng = np.random.default_rng(seed)
    big = 2 * size
    jobs = [(big, seed * 100003 + i, int(rng.integers(qualities[0], qualities[1] + 1))) for i in range(count)]
    n = SAMPLES_PER_CROP * count
    x0 = np.empty((n, size, size), np.float32)
    target = np.empty_like(x0)
    qs = np.empty((n, 64), np.float32)

    def quarters(p):
        return [p[:size, :size], p[:size, size:], p[size:, :size], p[size:, size:]]

    chunk = 64  # graphs of at most 64 images
    ctx = multiprocessing.get_context("spawn")
    with ctx.Pool(max(1, workers)) as pool:
        encoded = pool.imap(_encode_crop, jobs, chunksize=16)
        for c0 in range(0, count, chunk):
            coefs_list = [next(encoded) for _ in range(min(chunk, count - c0))]
            imgs = []
            for k, coefs in enumerate(coefs_list):
                i = SAMPLES_PER_CROP * (c0 + k)
                imgs.append(BatchImage(coefs, weight, [pweight] * 3, target_iterations, big, big, "tgv"))
                with Plan(coefs, weight, [pweight] * 3, 0, big, big) as plan:
                    plan.init()
                    y, cb, cr = (p.reshape(big, big) for p in plan.download())
                x0[i:i + 4] = quarters(y)
                # the component images: every 2x2 group of the upsampled start image is one value
                x0[i + 4] = cb[::2, ::2]
                x0[i + 5] = cr[::2, ::2]
                qs[i:i + 4] = quant_table(coefs[0])
                qs[i + 4] = quant_table(coefs[1])
                qs[i + 5] = quant_table(coefs[2])
            with Batch(imgs, streams=4, tune=False) as b:
                b.run()
                for k in range(len(imgs)):
                    i = SAMPLES_PER_CROP * (c0 + k)
                    y, cb, cr = (p.reshape(big, big).astype(np.float64) for p in b.download(k))
                    target[i:i + 4] = quarters(y)
                    for j, p in ((4, cb), (5, cr)):
                        target[i + j] = 0.25 * (p[0::2, 0::2] + p[1::2, 0::2] + p[0::2, 1::2] + p[1::2, 1::2])
            if log and ((c0 + chunk) % 1024 == 0 or c0 + chunk >= count):
                log(f"solved {min(c0 + chunk, count)} of {count} crops", flush=True)
    return x0, qs, target


def train(net: WarmStartNet, x0, q, target, steps: int, batch: int = 16, lr: float = 1e-3, seed: int = 0,
          log_every: int = 100, log=print):
    """Adam over random mini-batches of the resident set; cosine-decayed learning rate."""
    net.set_dataset(x0, q, target)
    rng = np.random.default_rng(seed)
    losses = []
    t0 = time.perf_counter()
    for s in range(steps):
        idx = rng.integers(0, x0.shape[0], size=batch)
        cur = lr * 0.5 * (1.0 + np.cos(np.pi * s / steps))
        losses.append(net.train_step(idx, cur))
        if log and (s % log_every == 0 or s == steps - 1):
            log(f"step {s} loss {losses[-1]:.6g} ({time.perf_counter() - t0:.1f} s)")
    return losses


@dataclass
class WarmResult:
    planes: list
    objective: float
    used_fallback: bool
    warm_objective: float
    warm_planes: list  # the refined prediction, also when the plain result is returned


def warm_solve(coefs, W, H, net: WarmStartNet, iterations: int, plain_objective: float | None = None,
               weight: float = 0.3, pweight=(0.001, 0.001, 0.001), plain_iterations: int = 100,
               step_iterations: int | None = None, step_from: int = 0, momentum_offset: int = 0,
               precision: str = "f32_fast") -> WarmResult:
    """CNN prediction -> projection -> ``iterations`` plain iterations with the step of the
    ``step_iterations`` schedule (default: ``plain_iterations``; with ``step_from`` the step decays
    from the ``step_from`` schedule's to it over the refinement) and the FISTA recurrence started
    ``momentum_offset`` steps in (default 0: as the reference starts it); if the result's
    objective is higher than the plain ``plain_iterations`` result's, the plain result is
    returned.

    The prediction is not bit-exact to anything (the CNN runs in bf16), so the refinement takes
    the fast f32 gradient sweep by default (``precision``); the plain result is the exact loop.
    Two settings measured at 2048x1536 4:2:0 q50 against plain 100 (6.3 ms): 15 iterations with
    the step decaying from the 3- to the 600-iteration schedule's and the recurrence started 10
    steps in take 22 % of its time and end 2.2 % above its objective (the fallback is taken); 23
    iterations with the step decaying from 2 to 1500 and the recurrence started 10 steps in reach
    its objective within 0.1 % in 28 % of its time."""
    with Plan(coefs, weight, list(pweight), iterations, W, H) as plan:
        plan.set_precision(precision)
        _lib.check(plan._lib.artefact_hip_plan_set_step_schedule(
            plan._h, step_from, plain_iterations if step_iterations is None else step_iterations), "step")
        _lib.check(plan._lib.artefact_hip_plan_set_momentum_offset(plan._h, momentum_offset), "momentum")
        _lib.check(plan._lib.artefact_hip_plan_warm_init(plan._h, net._h), "artefact_hip_plan_warm_init")
        plan.iterate()
        planes = plan.download()
        fw, _ = plan.objective()
    if plain_objective is None:
        with Plan(coefs, weight, list(pweight), plain_iterations, W, H) as plan:
            plan.init()
            plan.iterate()
            plain = plan.download()
            plain_objective, _ = plan.objective()
    else:
        plain = None
    if fw <= plain_objective:
        return WarmResult(planes, fw, False, fw, planes)
    if plain is None:
        with Plan(coefs, weight, list(pweight), plain_iterations, W, H) as plan:
            plan.init()
            plan.iterate()
            plain = plan.download()
    return WarmResult(plain, plain_objective, True, fw, planes)


def main(argv=None) -> int:
    """python -m artefact_amd.warmstart --crops N --steps S --out weights.npz"""
    import argparse

    ap = argparse.ArgumentParser(description="train the warm-start CNN on solver-generated pairs")
    ap.add_argument("--crops", type=int, default=8000, help="crops of 2 size x 2 size (six samples each)")
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--steps", type=int, default=60000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--target-iterations", type=int, default=300)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workers", type=int, default=8, help="processes that synthesise and encode the crops")
    ap.add_argument("--out", default=WEIGHTS_FILE)
    args = ap.parse_args(argv)
    t0 = time.perf_counter()
    x0, q, target = training_pairs(args.crops, args.size, args.seed, args.target_iterations, workers=args.workers)
    t_data = time.perf_counter() - t0
    print(f"pairs: {x0.shape[0]} planes of {args.size}x{args.size} in {t_data:.1f} s", flush=True)
    net = WarmStartNet(init_params(args.seed))
    t0 = time.perf_counter()
    losses = train(net, x0, q, target, args.steps, args.batch, args.lr, args.seed)
    t_train = time.perf_counter() - t0
    base = float(np.mean(((x0 - target) / 64.0) ** 2))
    print(f"trained {args.steps} steps in {t_train:.1f} s; final loss {np.mean(losses[-50:]):.6g} "
          f"(identity map: {base:.6g})", flush=True)
    net.save(args.out, crops=args.crops, size=args.size, steps=args.steps, batch=args.batch, lr=args.lr,
             target_iterations=args.target_iterations, seed=args.seed, data_seconds=t_data,
             train_seconds=t_train, final_loss=float(np.mean(losses[-50:])), identity_loss=base,
             input_features=INPUT_FEATURES)
    net.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
