"""Benchmark of the reconstruction loop on one MI355X.

Workload: plain TGV, YCbCr 4:2:0, 2048x1536 (Y 2048x1536, Cb/Cr 1024x768), IJG
tables at quality 75, 100 iterations, f32, synthetic seeded image.  One step is
one solve from resident coefficients: the start image (dequantise + IDCT +
upsample) and the 100-iteration loop (a hipGraph of 200 kernels after one that clears
the norm accumulators), inputs already
in HBM.  ``value`` is Mpixel*iterations/s over the timed steps.

The JSON line also carries
* ``roofline``: the iteration's algorithmic HBM bytes B = sum_c (20 N + 16 M_c)
  (x, x_prev, x_new as f32 over the full grid; int16 coefficients and the f32
  coefficient estimate read and written per component coefficient) divided by
  the measured mean duration of one iteration's two kernels (HIP events around
  each launch on the plan's stream), against 8 TB/s;
* ``cpu_baseline``: the CPU oracle (a restatement of the reference's scalar loop, with
  its threading: the FISTA update one thread per channel, the rest serial) timed on this
  host on a bounded sample of the same workload (same geometry, 30 iterations, loop only),
  with the single-threaded figure beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ITERATION_KERNELS = ("k_gradient", "k_project")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_tgv_420_2048x1536.json")


def summarize_pmc(read_csv, write_csv, copy_bytes=256 << 20):
    """Per-kernel L2-to-memory (EA) bytes per dispatch from two rocprofv3 --pmc CSVs, sized by
    request: reads TCC_EA0_RDREQ_32B/_64B/_128B_sum (32, 64, 128 bytes each; TCC_EA0_RDREQ_sum
    checks that they cover every request), writes TCC_EA0_WRREQ_sum and TCC_EA0_WRREQ_64B_sum
    (64-byte requests, the rest 32).  These are L2 misses: what the XCDs' L2s fetch from and
    write to the memory side (Infinity Cache and HBM).  The calibration copy of copy_bytes (the
    largest __amd_rocclr_copyBuffer dispatches) is reported as a check, not used as a factor."""
    import csv
    from collections import defaultdict

    def load(path):
        per = defaultdict(lambda: defaultdict(list))  # kernel -> counter -> values per dispatch
        for r in csv.DictReader(open(path)):
            name = r["Kernel_Name"]
            if "copyBuffer" in name:
                short = "copy"
            elif "k_gradient" in name:
                short = "k_gradient_fast" if "true, true" in name else "k_gradient"
            else:
                m = re.search(r"\b(k_\w+)", name)
                short = m.group(1) if m else "other"
            per[short][r["Counter_Name"]].append(float(r["Counter_Value"]))
        return per

    rd, wr = load(read_csv), load(write_csv)

    def mean(d, k, c):
        v = d.get(k, {}).get(c, [])
        return sum(v) / len(v) if v else 0.0

    def top_mean(d, c):  # the calibration copies: the largest copy dispatches of the run
        v = sorted(d.get("copy", {}).get(c, []))[-4:]
        return sum(v) / len(v) if v else 0.0

    def sized(get):
        n32, n64, n128 = (get("TCC_EA0_RDREQ_32B_sum"), get("TCC_EA0_RDREQ_64B_sum"),
                          get("TCC_EA0_RDREQ_128B_sum"))
        n, w, w64 = get("TCC_EA0_RDREQ_sum"), get("TCC_EA0_WRREQ_sum"), get("TCC_EA0_WRREQ_64B_sum")
        return {"read_requests": int(n), "read_unsized": int(n - n32 - n64 - n128),
                "read_bytes": int(32 * n32 + 64 * n64 + 128 * n128),
                "write_bytes": int(32 * (w - w64) + 64 * w64)}

    kernels = {}
    for k in sorted(set(rd) | set(wr)):
        if k == "copy":
            continue
        r = sized(lambda c: mean(rd, k, c) if c.startswith("TCC_EA0_RD") else mean(wr, k, c))
        r["bytes"] = r["read_bytes"] + r["write_bytes"]
        kernels[k] = r
    copy = sized(lambda c: top_mean(rd, c) if c.startswith("TCC_EA0_RD") else top_mean(wr, c))
    copy["expected_bytes_each_way"] = copy_bytes
    def total(names):
        return sum(kernels[k]["bytes"] for k in names if k in kernels)

    return {"counters": "L2-to-EA requests by size (TCC_EA0_RDREQ_32B/64B/128B, TCC_EA0_WRREQ/_64B)",
            "copy_check": copy, "kernels": kernels,
            "bytes_per_iteration": total(("k_gradient", "k_project_lane")),
            "bytes_per_iteration_fast": total(("k_gradient_fast", "k_project_lane"))}


def algorithmic_bytes(coefs, W, H):
    N = W * H
    return sum(20 * N + 16 * c.rounded_px_count for c in coefs)


def subsampling(coefs):
    if len(coefs) == 1:
        return "luma"
    f = (coefs[1].horizontal_samp_factor, coefs[1].vertical_samp_factor)
    return {(1, 1): "4:4:4", (2, 2): "4:2:0", (2, 1): "4:2:2", (1, 2): "4:4:0"}.get(f, str(f))


def describe(wl):
    kind = "plain TV" if wl["mode"] == "tv" else "plain TGV"
    chans = "luma only" if len(wl["coefs"]) == 1 else f"YCbCr {subsampling(wl['coefs'])}"
    return (f"{kind}, {chans}, {wl['width']}x{wl['height']}, IJG q{wl['quality']}, {wl['iterations']} iterations, "
            f"weight {wl['weight']}, pweight {wl['pweight'][0]}, seed {wl['seed']}")


def copy_calibration(nbytes=256 << 20, reps=4):
    """Device-to-device copy of a known size (HIP runtime copy kernel), timed with HIP
    events: the achievable streaming bandwidth on this box, and under rocprofv3 --pmc
    the FETCH_SIZE / WRITE_SIZE reference that scales the kernels' counters."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)):
        return None
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    hip.hipEventCreate(ctypes.byref(ev0))
    hip.hipEventCreate(ctypes.byref(ev1))
    hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)  # warm-up, hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipEventRecord(ev0, None)
    for _ in range(reps):
        hip.hipMemcpyAsync(dst, src, ctypes.c_size_t(nbytes), 3, None)
    hip.hipEventRecord(ev1, None)
    hip.hipEventSynchronize(ev1)
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), ev0, ev1)
    hip.hipEventDestroy(ev0)
    hip.hipEventDestroy(ev1)
    hip.hipFree(src)
    hip.hipFree(dst)
    return {"bytes_read_per_copy": nbytes, "copy_GBps": round(2 * nbytes * reps / (ms.value * 1e-3) / 1e9, 1)}


def fast_comparison(wl, args, exact_out, exact_objective, B):
    """The same workload with ARTEFACT_PRECISION_F32_FAST: throughput, iteration time and its
    roofline fraction, and the distance of its result from the bit-exact one."""
    import numpy as np

    import artefact_amd

    W, H, iters = wl["max_w"], wl["max_h"], wl["iterations"]
    plan = artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, mode=wl["mode"], graph=True)
    plan.set_precision("f32_fast")
    tuning = plan.tune()
    for _ in range(args.warmup):
        plan.init()
        plan.iterate()
    plan.download()
    plan.set_timing(args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.init()
        plan.iterate()
    plan.last_norms()
    elapsed = time.perf_counter() - t0
    loop_ms, launches = plan.timing()
    t_iter_ms = loop_ms / (launches * iters)
    plan.init()
    plan.iterate()
    objective, _ = plan.objective()
    out = plan.download()
    plan.close()
    diff = np.concatenate([np.abs(a - b) for a, b in zip(out, exact_out)])
    achieved = B / (t_iter_ms * 1e-3) / 1e9
    return {
        "precision": "f32_fast",
        "value": round(W * H * iters * args.steps / elapsed / 1e6, 2),
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "iteration_ms": round(t_iter_ms, 5),
        "roofline_frac": round(achieved / HBM_PEAK_GBS, 4),
        "objective_rel_to_exact": float(f"{(objective - exact_objective) / exact_objective:.3g}"),
        "max_abs_vs_exact": float(f"{diff.max():.3g}"),
        "share_within_0.05_of_exact": float(f"{np.mean(diff <= 0.05):.6f}"),
        "autotuned": tuning,
    }


def host_cpu():
    """(CPU description, CPUs this process may run on).  The description is the first
    processor's model name with its family, model, stepping and clock from /proc/cpuinfo (a
    virtual machine may report only a generic model name)."""
    fields = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break  # first processor only
                k, _, v = line.partition(":")
                fields[k.strip()] = v.strip()
    except OSError:
        pass
    model = fields.get("model name")
    extra = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping", "cpu MHz") if k in fields]
    if model and extra:
        model = f"{model} ({', '.join(extra)})"
    return model, len(os.sched_getaffinity(0))


def cpu_baseline(wl, iters, runs=3):
    """The oracle's loop on this host with the reference's own threading: one thread per channel
    for the FISTA update (pipeline_scalar/mod.rs:54-59, rayon), compute_step serial
    (compute_step.rs:19-71).  The single-threaded figure is reported beside it."""
    import oracle

    W, H = wl["max_w"], wl["max_h"]
    C = len(wl["coefs"])

    def best(threads):
        return min(oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, return_time=True,
                                  threads=threads)[1] for _ in range(runs))

    secs = best(C)
    secs1 = best(1) if C > 1 else secs
    model, ncpu = host_cpu()
    return {
        "value": W * H * iters / secs / 1e6,
        "unit": "Mpixel*it/s",
        "cores": C,
        "kind": "port",
        "single_thread_value": W * H * iters / secs1 / 1e6,
        "cpu_model": model,
        "host_cpus": ncpu,
        "sample": f"{describe(wl)} geometry, {iters} iterations, loop only, best of {runs} runs ({secs:.2f} s; "
                  f"{secs1:.2f} s on one thread), C restatement of pipeline_scalar (oracle/oracle.c, -O2 "
                  f"-ffp-contract=off) with the reference's threading: the FISTA update on {C} thread(s), one per "
                  f"channel (mod.rs:54-59), the prob, TV, second-order, norm, descent and projection steps serial "
                  f"(compute_step.rs:19-71)",
    }


def measure_plan(wl, precision="f32", steps=5, warmup=2):
    """One workload solved steps times from resident coefficients (init + graph of the loop) after
    autotuning: per-solve and per-iteration time (HIP events around the graph launches), the
    roofline fraction of the iteration, the objective and the planes."""
    import artefact_amd

    W, H, iters = wl["max_w"], wl["max_h"], wl["iterations"]
    plan = artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, mode=wl["mode"], graph=True)
    plan.set_precision(precision)
    tuning = plan.tune()
    for _ in range(warmup):
        plan.init()
        plan.iterate()
    plan.download()
    plan.set_timing(steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.init()
        plan.iterate()
    plan.last_norms()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    loop_ms, launches = plan.timing()
    plan.set_timing(0)
    plan.init()
    plan.iterate()
    objective, _ = plan.objective()
    out = plan.download()
    plan.close()
    t_iter = loop_ms / (launches * iters)
    B = algorithmic_bytes(wl["coefs"], W, H)
    achieved = B / (t_iter * 1e-3) / 1e9
    return {"ms_per_solve": round(ms, 4), "iteration_ms": round(t_iter, 5),
            "Mpixel_it_s": round(W * H * iters / (ms * 1e-3) / 1e6, 2),
            "roofline": {"achieved_GBps": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes_per_iteration": B},
            "objective": objective, "autotuned": tuning}, out


def measure_batch(imgs, steps=5, warmup=2, tune=True):
    """The images in one graph (stream count autotuned) against the same images solved one after
    another, each as its own graph-captured plan."""
    import artefact_amd

    work = sum(w["max_w"] * w["max_h"] * w["iterations"] for w in imgs)
    B = sum(algorithmic_bytes(w["coefs"], w["max_w"], w["max_h"]) * w["iterations"] for w in imgs)
    images = [artefact_amd.BatchImage(w["coefs"], w["weight"], w["pweight"], w["iterations"], w["max_w"], w["max_h"],
                                      w["mode"]) for w in imgs]
    t0 = time.perf_counter()
    batch = artefact_amd.Batch(images, streams=0, tune=tune)
    setup_s = time.perf_counter() - t0
    streams = batch.streams
    for _ in range(warmup):
        batch.run()
    batch.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
    batch.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    batch.close()
    plans = [artefact_amd.Plan(w["coefs"], w["weight"], w["pweight"], w["iterations"], w["max_w"], w["max_h"],
                               mode=w["mode"]) for w in imgs]
    for pl in plans:
        pl.init()
        pl.iterate()
    for pl in plans:
        pl.last_norms()
    reps = max(1, steps // 2)
    t0 = time.perf_counter()
    for _ in range(reps):
        for pl in plans:
            pl.init()
            pl.iterate()
            pl.last_norms()
    seq_ms = 1e3 * (time.perf_counter() - t0) / reps
    for pl in plans:
        pl.close()
    achieved = B / (ms * 1e-3) / 1e9
    return {"images": len(imgs), "streams": streams, "pixel_iterations": work, "ms_per_batch": round(ms, 3),
            "Mpixel_it_s": round(work / (ms * 1e-3) / 1e6, 2), "sequential_ms": round(seq_ms, 3),
            "batch_speedup_vs_sequential": round(seq_ms / ms, 3), "setup_s": round(setup_s, 2),
            "roofline": {"achieved_GBps": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes_per_batch": B}}


def measure_warm(net, coefs, W, H, K=15, steps=5, warmup=2):
    """CNN prediction -> projection -> K plain iterations (WARM15: the step decaying from the
    3- to the 600-iteration schedule's, the FISTA recurrence started 10 steps in) against the
    plain 100-iteration solve: times, objectives, the fallback decision, and the
    planes (decoded, plain 100, warm, plain K)."""
    import artefact_amd

    pw = [0.001] * 3

    def timed(plan, start):
        for _ in range(warmup):
            start(plan)
            plan.iterate()
        plan.last_norms()
        t0 = time.perf_counter()
        for _ in range(steps):
            start(plan)
            plan.iterate()
        plan.last_norms()
        return 1e3 * (time.perf_counter() - t0) / steps

    def warm_start(p):
        # The warm start is not bit-exact to anything (the CNN runs in bf16), so its refinement
        # takes the fast f32 gradient sweep.
        p.set_precision("f32_fast")
        artefact_amd._lib.check(p._lib.artefact_hip_plan_warm_init(p._h, net._h), "artefact_hip_plan_warm_init")

    def cnn_only(p):  # the prediction and its projection, no iteration
        warm_start(p)

    res, planes = {}, {}
    plain = artefact_amd.Plan(coefs, 0.3, pw, 100, W, H)
    res["plain_100_ms"] = round(timed(plain, lambda p: p.init()), 4)
    plain.init()
    plain.iterate()
    res["plain_100_objective"], _ = plain.objective()
    planes["plain_100"] = plain.download()
    plain.init()
    planes["decoded"] = plain.download()
    plain.close()
    warm = artefact_amd.Plan(coefs, 0.3, pw, K, W, H)
    set_refinement(warm, *WARM15)
    res[f"cnn_plus_{K}_ms"] = round(timed(warm, warm_start), 4)
    warm_start(warm)
    warm.iterate()
    res[f"cnn_plus_{K}_objective"], _ = warm.objective()
    planes[f"cnn_plus_{K}"] = warm.download()
    warm.close()
    zero = artefact_amd.Plan(coefs, 0.3, pw, 0, W, H)
    for _ in range(warmup):
        cnn_only(zero)
    zero.last_norms()
    t0 = time.perf_counter()
    for _ in range(steps):
        cnn_only(zero)
    zero.last_norms()
    res["cnn_and_projection_ms"] = round(1e3 * (time.perf_counter() - t0) / steps, 4)
    zero.close()
    cold = artefact_amd.Plan(coefs, 0.3, pw, K, W, H)
    res[f"plain_{K}_ms"] = round(timed(cold, lambda p: p.init()), 4)
    cold.init()
    cold.iterate()
    res[f"plain_{K}_objective"], _ = cold.objective()
    planes[f"plain_{K}"] = cold.download()
    cold.close()
    f_plain, f_warm = res["plain_100_objective"], res[f"cnn_plus_{K}_objective"]
    res["time_to_plain_100_objective"] = warm_time_to_objective(net, coefs, W, H, f_plain, res["plain_100_ms"],
                                                                 timed, warm_start)
    res["warm_time_fraction_of_plain"] = round(res[f"cnn_plus_{K}_ms"] / res["plain_100_ms"], 4)
    res["warm_objective_rel_to_plain"] = float(f"{(f_warm - f_plain) / f_plain:.4g}")
    res["reaches_plain_objective_within_0.1pct"] = bool(f_warm <= f_plain * 1.001)
    res["used_fallback"] = bool(f_warm > f_plain)
    return res, planes


# Warm refinement settings (S1, S2, T): the step decays from the S1-iteration schedule's to the
# S2-iteration schedule's over the refinement (S1 = 0: the constant step of S2), and the FISTA
# recurrence starts T steps in.  CNN + 15 takes WARM15; the time-to-objective search tries
# WARM_SETTINGS.
WARM15 = (3, 600, 10)
WARM_SETTINGS = ((0, 400, 60), (50, 600, 40), (15, 600, 20), (6, 1500, 20), (3, 600, 20), (3, 1500, 10),
                 (2, 1500, 10), (1, 1500, 10))


def set_refinement(p, S1, S2, T):
    import artefact_amd

    artefact_amd._lib.check(p._lib.artefact_hip_plan_set_step_schedule(p._h, S1, S2), "step schedule")
    artefact_amd._lib.check(p._lib.artefact_hip_plan_set_momentum_offset(p._h, T), "momentum")


def warm_time_to_objective(net, coefs, W, H, f_plain, plain_ms, timed, warm_start, settings=WARM_SETTINGS,
                           tolerance=1e-3):
    """The fewest plain iterations K after the CNN start whose objective reaches plain 100's
    within `tolerance` (the target's 0.1 %), for refinement steps decaying from an S1- to an
    S2-iteration schedule's and
    the FISTA recurrence started T steps in: for each (S1, S2, T) of `settings` (see WARM15) a scan
    of K = 15, 20, .., 100, then one by one below the first K that reaches (with a late start the
    objective need not fall monotonically with K), and the fastest setting timed like the other
    solves."""
    import artefact_amd

    target = f_plain * (1.0 + tolerance)

    pw = [0.001] * 3

    def objective(K, S1, S2, T):
        with artefact_amd.Plan(coefs, 0.3, pw, K, W, H) as p:
            set_refinement(p, S1, S2, T)
            warm_start(p)
            p.iterate()
            return p.objective()[0]

    best = None
    for S1, S2, T in settings:
        hit = next((K for K in range(15, 101, 5) if objective(K, S1, S2, T) <= target), None)
        if hit is None:
            continue
        hit = next((K for K in range(max(15, hit - 4), hit) if objective(K, S1, S2, T) <= target), hit)
        if best is None or hit < best[0]:
            best = (hit, S1, S2, T)
    criterion = f"objective <= plain 100's x {1.0 + tolerance:g}"
    if best is None:
        return {"reached": False, "criterion": criterion,
                "searched": "K in 15..100 for (step schedule, momentum offset) in %s" % list(settings)}
    K, S1, S2, T = best
    with artefact_amd.Plan(coefs, 0.3, pw, K, W, H) as p:
        set_refinement(p, S1, S2, T)
        ms = timed(p, warm_start)
        warm_start(p)
        p.iterate()
        f = p.objective()[0]
    return {"reached": True, "criterion": criterion, "cnn_plus_K": K, "step_from_iterations": S1,
            "step_iterations": S2, "momentum_offset": T,
            "ms": round(ms, 4),
            "objective_rel_to_plain": float(f"{(f - f_plain) / f_plain:.4g}"),
            "fraction_of_plain_time": round(ms / plain_ms, 4)}


GOLDEN = os.path.join(ROOT, "tests", "golden")


def golden(name):
    """The oracle's committed results for a full-size workload (tests/golden/make_config_fixtures.py)."""
    path = os.path.join(GOLDEN, "config_results.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(name)


def subsample_psnr(planes, fx, W, H):
    """PSNR (0..255 scale) of planes against a committed FP64 subsample."""
    import numpy as np

    sub = fx["fp64_subsample"]
    with np.load(os.path.join(GOLDEN, sub["file"]), allow_pickle=False) as z:
        ref = [z[f"p{c}"].astype(np.float64) for c in range(len(planes))]
    s, o = sub["step"], sub["offset"]
    mse = np.mean([np.mean((p.reshape(H, W)[o::s, o::s].astype(np.float64) - r) ** 2) for p, r in zip(planes, ref)])
    return round(float(10 * np.log10(255.0 ** 2 / mse)), 3)


def prepare_configs(args):
    """The other configurations' inputs, synthesised before any GPU work (the batch's 64 images by a
    pool of spawned processes)."""
    from artefact_amd import synth

    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    t0 = time.perf_counter()
    pre = {
        "tv_luma_512": synth.workload("tv_luma_512", seed=0),
        "tgv_444_4096": synth.workload("tgv_444_4096", seed=0),
        "batch64": synth.batch_workload(64, seed=0, workers=workers),
        "warm": synth.encode(synth.synth_rgb(1536, 2048, 0), subsampling="420", quality=50),
    }
    print(f"bench: other configurations synthesised in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    return pre


def other_configs(args, pre):
    """The other BASELINE configurations, timed on this GPU in the same run as the headline (a few
    solves each): plain TV luma 512x512; the precision form at 4096x3072 4:4:4 q30 (f32 and
    BF16x3); the 64-image batch; the learned warm start at 2048x1536 4:2:0 q50."""
    from artefact_amd import warmstart

    steps, warmup = max(2, min(args.steps, 10)), max(1, min(args.warmup, 2))
    out = {}
    wl = pre["tv_luma_512"]
    r, _ = measure_plan(wl, "f32", steps, warmup)
    out["tv_luma_512"] = {"workload": describe(wl), **r}
    print(f"bench: tv_luma_512 {r['iteration_ms'] * 1e3:.2f} us per iteration", file=sys.stderr, flush=True)

    wl = pre["tgv_444_4096"]
    fx = golden("tgv_444_4096")
    prec = {}
    for p in ("f32", "bf16x3"):
        r, planes = measure_plan(wl, p, max(2, steps // 2), warmup)
        if fx:
            r["objective_rel_to_fp64_oracle"] = float(f"{(r['objective'] - fx['fp64_objective']) / fx['fp64_objective']:.4g}")
            r["psnr_vs_fp64_oracle_subsample_db"] = subsample_psnr(planes, fx, wl["max_w"], wl["max_h"])
        prec[p] = r
        print(f"bench: tgv_444_4096 {p} {r['ms_per_solve']:.2f} ms per solve", file=sys.stderr, flush=True)
    out["precision_444_4096"] = {"workload": describe(wl) + "; bf16x3: projection DCT/IDCTs on BF16 MFMA, hi/lo "
                                 "split, f32 accumulation", "by_precision": prec,
                                 "fp64_reference": "tests/golden/config_results.json (oracle.compute_f64, 100 "
                                                   "iterations; subsample every 16th row and column)" if fx else None}

    b = measure_batch(pre["batch64"], max(2, steps // 2), warmup)
    out["batch64"] = {"workload": "64 images, sides uniform in 256..2048 (multiples of 16), half 4:2:0 / half 4:4:4, "
                                  "IJG q30..90, half TV / half TGV, 100 iterations, one hipGraph over autotuned streams, "
                                  "autotuned per geometry, seed 0", **b}
    print(f"bench: batch64 {b['ms_per_batch']:.1f} ms per batch", file=sys.stderr, flush=True)

    try:
        net = warmstart.WarmStartNet.load()
    except (OSError, ValueError) as e:
        out["warm_start_420_2048_q50"] = {"error": str(e)}
        return out
    coefs = pre["warm"]
    W, H = coefs[0].rounded_px_w, coefs[0].rounded_px_h
    w, planes = measure_warm(net, coefs, W, H, 15, steps, warmup)
    net.close()
    fx = golden("warm_420_2048_q50")
    if fx:
        w["psnr_vs_fp64_converged_subsample_db"] = {k: subsample_psnr(v, fx, W, H) for k, v in planes.items()}
        w["fp64_converged_reference"] = ("tests/golden/config_results.json (oracle.compute_f64, 300 iterations; "
                                         "subsample every 8th row and column)")
    out["warm_start_420_2048_q50"] = {"workload": "learned warm start: TGV 4:2:0 2048x1536 IJG q50, CNN + 15 plain "
                                                  "iterations (step decaying from the 3- to the 600-iteration "
                                                  "schedule's, FISTA recurrence started 10 steps in, fast f32 "
                                                  "gradient sweep) vs plain 100 (exact)", **w}
    print(f"bench: warm start {w['cnn_plus_15_ms']:.3f} ms vs plain {w['plain_100_ms']:.3f} ms", file=sys.stderr,
          flush=True)
    return out


def bench_precision(args) -> int:
    """Precision form on the 4:4:4 4096x3072 q30 TGV workload: the projection's transforms on
    BF16 MFMA (hi/lo, f32 accumulation); error of the result against the FP64 oracle loop per
    channel, next to the plain f32 path's."""
    import numpy as np

    import artefact_amd
    import oracle
    from artefact_amd import synth

    wl = synth.workload("tgv_444_4096", seed=0)
    W, H, iters = wl["max_w"], wl["max_h"], wl["iterations"]
    coefs = wl["coefs"]
    print(f"bench: precision form {W}x{H} iters={iters}", file=sys.stderr, flush=True)
    res_by = {}
    outs = {}
    for prec in ("f32", "bf16x3"):
        plan = artefact_amd.Plan(coefs, wl["weight"], wl["pweight"], iters, W, H, mode=wl["mode"], graph=True)
        plan.set_precision(prec)
        tuned = plan.tune()
        for _ in range(args.warmup):
            plan.init()
            plan.iterate()
        plan.download()
        plan.set_timing(args.steps)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            plan.init()
            plan.iterate()
        plan.last_norms()
        ms = 1e3 * (time.perf_counter() - t0) / args.steps
        loop_ms, launches = plan.timing()
        plan.set_timing(0)
        plan.init()
        prof = plan.profile()
        plan.init()
        plan.iterate()
        obj, _ = plan.objective()
        outs[prec] = plan.download()
        plan.close()
        res_by[prec] = {"ms_per_step": round(ms, 4), "iteration_ms": round(loop_ms / (launches * iters), 5),
                        "eager_breakdown_ms": {"k_gradient": round(prof["k_gradient_ms"], 5),
                                               "k_project": round(prof["k_project_ms"], 5)},
                        "objective": obj, "autotuned": tuned}
        print(f"bench: {prec} {ms:.3f} ms/solve", file=sys.stderr, flush=True)
    errors = None
    if not args.no_error:
        t0 = time.perf_counter()
        ref = oracle.compute_f64(coefs, wl["weight"], wl["pweight"], iters, W, H)
        fref, _ = oracle.objective(coefs, wl["weight"], wl["pweight"], W, H, ref)
        errors = {"fp64_oracle_s": round(time.perf_counter() - t0, 1)}
        for prec, out in outs.items():
            e = [np.abs(o.astype(np.float64) - r) for o, r in zip(out, ref)]
            errors[prec] = {
                "max_abs_per_channel": [round(float(x.max()), 5) for x in e],
                "rms_per_channel": [float(f"{np.sqrt((x ** 2).mean()):.4g}") for x in e],
                "rel_objective_error": float(f"{(res_by[prec]['objective'] - fref) / fref:.3g}"),
            }
    ms = res_by["bf16x3"]["ms_per_step"]
    B = algorithmic_bytes(coefs, W, H)
    t_iter_ms = res_by["bf16x3"]["iteration_ms"]
    achieved = B / (t_iter_ms * 1e-3) / 1e9
    res = {
        "metric": "Mpixel*iterations/s + time to plain-100-iter objective",
        "value": round(W * H * iters / (ms * 1e-3) / 1e6, 2),
        "unit": "Mpixel*it/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3 transforms, f32 stencil, f64 objective",
        "data": "synthetic (seeded generator, artefact_amd/synth.py)",
        "config": {
            "workload": "precision form: " + describe(wl) + "; projection DCT/IDCTs on BF16 MFMA with hi/lo split",
            "by_precision": res_by,
            "error_vs_fp64_oracle": errors,
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "one iteration = k_gradient + k_project_mfma (HIP events around the "
                      "timed region's graph launches, divided by the iterations)",
            "algorithmic_bytes_per_iteration": B,
        },
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_warmstart(args) -> int:
    """Learned warm start on 2048x1536 4:2:0 IJG q50 TGV: CNN prediction -> projection -> K plain
    iterations, against the plain 100-iteration loop (time and objective), with the plain
    fallback, and PSNR of each result against the FP64 loop run to convergence."""
    import numpy as np

    import artefact_amd
    import oracle
    from artefact_amd import synth, warmstart

    W, H, K = 2048, 1536, args.warm_iterations
    coefs = synth.encode(synth.synth_rgb(H, W, 0), subsampling="420", quality=50)
    pw = [0.001] * 3
    net = warmstart.WarmStartNet.load(args.weights)
    meta = dict(np.load(args.weights, allow_pickle=False))
    print(f"bench: warm start {W}x{H} q50, K={K}", file=sys.stderr, flush=True)

    def timed(plan, start):
        for _ in range(args.warmup):
            start(plan)
            plan.iterate()
        plan.last_norms()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            start(plan)
            plan.iterate()
        plan.last_norms()
        return 1e3 * (time.perf_counter() - t0) / args.steps

    plain = artefact_amd.Plan(coefs, 0.3, pw, 100, W, H)
    t_plain = timed(plain, lambda p: p.init())
    plain.init()
    plain.iterate()
    f_plain, _ = plain.objective()
    x_plain = plain.download()
    plain.close()

    def warm_start(p):
        # The warm start is not bit-exact to anything (the CNN runs in bf16), so its refinement
        # takes the fast f32 gradient sweep.
        p.set_precision("f32_fast")
        artefact_amd._lib.check(p._lib.artefact_hip_plan_warm_init(p._h, net._h), "artefact_hip_plan_warm_init")

    warm = artefact_amd.Plan(coefs, 0.3, pw, K, W, H)
    set_refinement(warm, *WARM15)
    t_warm = timed(warm, warm_start)
    warm_start(warm)
    warm.iterate()
    f_warm, _ = warm.objective()
    x_warm = warm.download()
    warm.close()

    cold = artefact_amd.Plan(coefs, 0.3, pw, K, W, H)
    t_cold = timed(cold, lambda p: p.init())
    cold.init()
    cold.iterate()
    f_cold, _ = cold.objective()
    x_cold = cold.download()
    cold.init()
    x_dec = cold.download()
    cold.close()
    sweep = {}
    for k in args.warm_sweep:
        for kind in ("warm", "plain", "warm_step100"):
            pl = artefact_amd.Plan(coefs, 0.3, pw, k, W, H)
            if kind == "warm_step100":
                artefact_amd._lib.check(pl._lib.artefact_hip_plan_set_step_iterations(pl._h, 100), "step")
            t_k = timed(pl, warm_start if kind.startswith("warm") else (lambda q: q.init()))
            (warm_start(pl) if kind.startswith("warm") else pl.init())
            pl.iterate()
            f_k, _ = pl.objective()
            pl.close()
            sweep[f"{kind}_{k}"] = {"ms": round(t_k, 4), "objective_rel_to_plain_100": float(f"{(f_k - f_plain) / f_plain:.4g}")}
    fallback = f_warm > f_plain
    reach = warm_time_to_objective(net, coefs, W, H, f_plain, t_plain, timed, warm_start)
    print(f"bench: plain {t_plain:.3f} ms F={f_plain:.6g}; warm {t_warm:.3f} ms F={f_warm:.6g}; "
          f"plain {K} {t_cold:.3f} ms F={f_cold:.6g}", file=sys.stderr, flush=True)
    psnr = None
    if not args.no_error:
        print(f"bench: FP64 oracle, {args.converged_iterations} iterations", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        ref = oracle.compute_f64(coefs, 0.3, pw, args.converged_iterations, W, H)

        def p(x):
            mse = np.mean([np.mean((a.astype(np.float64) - r) ** 2) for a, r in zip(x, ref)])
            return round(float(10 * np.log10(255.0 ** 2 / mse)), 3)

        psnr = {"decoded": p(x_dec), "plain_100": p(x_plain), f"warm_cnn_plus_{K}": p(x_warm),
                f"plain_{K}": p(x_cold), "fp64_seconds": round(time.perf_counter() - t0, 1),
                "fp64_iterations": args.converged_iterations}
    work = W * H * 100
    res = {
        "metric": "Mpixel*iterations/s + time to plain-100-iter objective",
        "value": round(work / ((t_warm + (t_plain if fallback else 0.0)) * 1e-3) / 1e6, 2),
        "unit": "Mpixel*it/s (plain-100-equivalent work over the warm-started time)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_warm + (t_plain if fallback else 0.0), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 loop, bf16 CNN (f32 accumulation)",
        "data": "synthetic (seeded generator, artefact_amd/synth.py)",
        "config": {
            "workload": f"learned warm start: TGV 4:2:0 {W}x{H} IJG q50, CNN + {K} plain iterations (step of the "
                        f"100-iteration schedule, fast f32 gradient sweep) vs plain 100 (exact)",
            "plain_100_ms": round(t_plain, 4), "plain_100_objective": f_plain,
            f"warm_ms": round(t_warm, 4), "warm_objective": f_warm,
            "warm_time_fraction_of_plain": round(t_warm / t_plain, 4),
            "warm_objective_rel_to_plain": float(f"{(f_warm - f_plain) / f_plain:.4g}"),
            "reaches_plain_objective_within_0.1pct": bool(f_warm <= f_plain * 1.001),
            "used_fallback": bool(fallback),
            f"plain_{K}_ms": round(t_cold, 4), f"plain_{K}_objective": f_cold,
            "psnr_vs_fp64_converged_db": psnr,
            "iteration_sweep": sweep,
            "time_to_plain_100_objective": reach,
            "training": {k: v.item() if hasattr(v, "item") and np.asarray(v).ndim == 0 else None
                         for k, v in meta.items() if k != "params"},
        },
    }
    net.close()
    print(json.dumps(res), flush=True)
    return 0


def bench_process(args) -> int:
    """Artefact::process (lib.rs:105-211) host to host: a 2048x1536 4:2:0 q75 JPEG (Pillow-encoded
    synthetic image) decoded on the host, then the loop and the output stage on the GPU
    (artefact_hip_process: plan set-up, upload, 100 iterations, RGB, download, release).  The CPU
    baseline is the oracle's whole pipeline on the same coefficients with the f64 norm, so the
    two RGB images must be identical."""
    import io

    import numpy as np
    from PIL import Image

    import artefact_amd
    import oracle
    from artefact_amd import jpeg as jpeg_mod
    from artefact_amd import synth
    from artefact_amd.process import Artefact, ForAll, JpegSource, process_coefficients

    W, H, iters = 2048, 1536, 100
    rgb = np.stack([np.asarray(p, dtype=np.uint8) for p in synth.synth_rgb(H, W, 0)], axis=-1)
    buf = io.BytesIO()
    Image.fromarray(rgb, "RGB").save(buf, "JPEG", quality=75, subsampling=2)  # 2: 4:2:0
    data = buf.getvalue()
    weight, pweight, its = [0.3] * 3, [0.001] * 3, [iters] * 3
    print(f"bench: process {W}x{H} 4:2:0 q75 JPEG ({len(data)} bytes), {iters} iterations",
          file=sys.stderr, flush=True)

    jp = jpeg_mod.decode(data)
    reps = max(args.steps, 5)
    t0 = time.perf_counter()
    for _ in range(reps):
        jpeg_mod.decode(data)
    decode_ms = 1e3 * (time.perf_counter() - t0) / reps

    for _ in range(args.warmup):
        process_coefficients(jp, weight, pweight, its)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = process_coefficients(jp, weight, pweight, its)
    gpu_ms = 1e3 * (time.perf_counter() - t0) / args.steps

    def from_bytes():
        return Artefact().source(JpegSource.Buffer(data)).weight(ForAll(0.3)).pweight(ForAll(0.001)) \
            .iterations(ForAll(iters)).process()

    from_bytes()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out_b = from_bytes()
    e2e_ms = 1e3 * (time.perf_counter() - t0) / args.steps

    # The loop's iteration as the product runs it (same geometry, events around graph launches).
    Wg = max(c.rounded_px_w for c in jp.coefs)
    Hg = max(c.rounded_px_h for c in jp.coefs)
    plan = artefact_amd.Plan(jp.coefs, 0.3, pweight, iters, Wg, Hg, graph=True)
    plan.init()
    plan.iterate()
    plan.set_timing(args.steps)
    for _ in range(args.steps):
        plan.init()
        plan.iterate()
    loop_ms, launches = plan.timing()
    plan.close()
    t_iter_ms = loop_ms / (launches * iters)
    B = algorithmic_bytes(jp.coefs, Wg, Hg)

    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        ref = oracle.process(jp.coefs, jp.real_px_w, jp.real_px_h, weight, pweight, its, norm_mode="f64")
        secs = time.perf_counter() - t0
        cpu = {"value": W * H * iters / secs / 1e6, "unit": "Mpixel*it/s", "cores": 1, "kind": "port",
               "sample": f"the whole pipeline after decoding ({iters} iterations, init, loop and output stage; "
                         f"{secs:.1f} s), single-threaded C restatement (oracle/oracle.c), f64 norm",
               "rgb_identical_to_gpu": bool(np.array_equal(ref, out) and np.array_equal(ref, out_b))}
    res = {
        "metric": "Mpixel*iterations/s + time to plain-100-iter objective",
        "value": round(W * H * iters / (e2e_ms * 1e-3) / 1e6, 2),
        "unit": "Mpixel*it/s (host to host: JPEG bytes in, RGB bytes out, PCIe included)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(e2e_ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator, artefact_amd/synth.py), JPEG-encoded by Pillow",
        "config": {
            "workload": f"Artefact::process: {W}x{H} YCbCr 4:2:0 IJG q75 JPEG, plain TGV, {iters} iterations, "
                        "weight 0.3, pweight 0.001, seed 0",
            "jpeg_bytes": len(data),
            "decode_ms": round(decode_ms, 3),
            "gpu_process_ms": round(gpu_ms, 3),
            "end_to_end_ms": round(e2e_ms, 3),
        },
        "roofline": {
            "bound": "hbm", "achieved": round(B / (t_iter_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(B / (t_iter_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "one loop iteration at this geometry (HIP events around graph launches)",
            "iteration_ms": round(t_iter_ms, 5), "algorithmic_bytes_per_iteration": B,
        },
    }
    if cpu is not None:
        res["cpu_baseline"] = cpu
    print(json.dumps(res), flush=True)
    return 0


def bench_batch(args) -> int:
    """The batched form: ``batchN`` = N images (synth.batch_workload) in one graph."""
    import artefact_amd
    from artefact_amd import synth

    count = int(args.workload[len("batch"):] or 64)
    imgs = synth.batch_workload(count, seed=0)
    print(f"bench: {count}-image batch generated", file=sys.stderr, flush=True)
    work = sum(w["max_w"] * w["max_h"] * w["iterations"] for w in imgs)
    B = sum(algorithmic_bytes(w["coefs"], w["max_w"], w["max_h"]) * w["iterations"] for w in imgs)
    images = [artefact_amd.BatchImage(w["coefs"], w["weight"], w["pweight"], w["iterations"], w["max_w"], w["max_h"],
                                      w["mode"]) for w in imgs]
    t0 = time.perf_counter()
    batch = artefact_amd.Batch(images, streams=0, tune=True)
    setup_s = time.perf_counter() - t0
    streams = batch.streams
    for _ in range(args.warmup):
        batch.run()
    batch.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    batch.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = work * args.steps / elapsed / 1e6
    batch.close()
    # The same images one after another, each as its own graph-captured plan.
    plans = [artefact_amd.Plan(w["coefs"], w["weight"], w["pweight"], w["iterations"], w["max_w"], w["max_h"],
                               mode=w["mode"]) for w in imgs]
    for pl in plans:
        pl.init()
        pl.iterate()
    for pl in plans:
        pl.last_norms()
    reps = max(1, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        for pl in plans:
            pl.init()
            pl.iterate()
            pl.last_norms()
    seq_ms = 1e3 * (time.perf_counter() - t0) / reps
    for pl in plans:
        pl.close()
    achieved = B / (ms_per_step * 1e-3) / 1e9
    res = {
        "metric": "Mpixel*iterations/s + time to plain-100-iter objective",
        "value": round(value, 2),
        "unit": "Mpixel*it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, artefact_amd/synth.py batch_workload)",
        "config": {
            "workload": f"batched form: {count} images, sides uniform in 256..2048 (multiples of 16), half 4:2:0 / "
                        f"half 4:4:4, IJG quality 30..90, half TV / half TGV, 100 iterations, one hipGraph over "
                        f"{streams} streams (count autotuned), autotuned per geometry, seed 0",
            "images": count,
            "pixel_iterations": work,
            "sequential_ms_per_step": round(seq_ms, 3),
            "batch_speedup_vs_sequential": round(seq_ms / ms_per_step, 3),
            "setup_s": round(setup_s, 2),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "whole batch: algorithmic bytes of every image's iterations over the step time",
            "algorithmic_bytes_per_step": B,
        },
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(imgs[0], args.cpu_iters)
    print(json.dumps(res), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="tgv_420_2048")
    ap.add_argument("--cpu-iters", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true", help="precision / warmstart: skip the FP64 oracle run")
    ap.add_argument("--weights", default=os.path.join(ROOT, "artefact_amd", "warmstart_weights.npz"))
    ap.add_argument("--warm-iterations", type=int, default=15)
    ap.add_argument("--converged-iterations", type=int, default=300)
    ap.add_argument("--warm-sweep", type=int, nargs="*", default=[15, 25, 40, 60])
    ap.add_argument("--copy-sweep", action="store_true", help="also time 16/64/128 MiB copies")
    ap.add_argument("--precision", default="f32", choices=["f32", "f32_fast"],
                    help="arithmetic of the timed loop (default: the bit-exact f32 drop-in)")
    ap.add_argument("--no-fast", action="store_true", help="skip the f32_fast comparison of the default line")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the other configurations (512x512 TV, precision form, batch, warm start) of the "
                         "default line")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("READ_CSV", "WRITE_CSV"),
                    help="print per-kernel L2-to-memory bytes from two rocprofv3 --pmc CSVs (request-size "
                         "counters) and exit")
    args = ap.parse_args()
    if args.summarize_pmc:
        print(json.dumps(summarize_pmc(*args.summarize_pmc), indent=1))
        return 0
    if args.gpus != 1:
        print(f"bench.py: the reconstruction loop runs on one GPU; --gpus {args.gpus} is not supported "
              f"(sharding images or plane bands across GPUs is not implemented)", file=sys.stderr)
        return 2

    import numpy as np

    import artefact_amd
    from artefact_amd import synth

    if args.workload.startswith("batch"):
        return bench_batch(args)
    if args.workload == "precision":
        return bench_precision(args)
    if args.workload == "warmstart":
        return bench_warmstart(args)
    if args.workload == "process":
        return bench_process(args)

    # Inputs of the other configurations first: the batch is synthesised by spawned processes,
    # which must start before this process touches the GPU.
    extras = args.workload == "tgv_420_2048" and not args.no_extras
    pre = prepare_configs(args) if extras else None
    wl = synth.workload(args.workload, seed=0)
    W, H, iters = wl["max_w"], wl["max_h"], wl["iterations"]
    coefs = wl["coefs"]
    C = len(coefs)
    print(f"bench: {args.workload} {W}x{H} C={C} iters={iters} steps={args.steps} warmup={args.warmup}",
          file=sys.stderr, flush=True)

    plan = artefact_amd.Plan(coefs, wl["weight"], wl["pweight"], iters, W, H, mode=wl["mode"], graph=True)
    plan.set_precision(args.precision)
    tuning = plan.tune()
    for _ in range(args.warmup):
        plan.init()
        plan.iterate()
    plan.download()  # synchronises

    # HIP events on the plan's stream around every iteration-graph launch of the timed region.
    plan.set_timing(args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.init()
        plan.iterate()
    plan.last_norms()  # synchronises the plan's stream
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = W * H * iters * args.steps / elapsed / 1e6
    loop_ms, launches = plan.timing()
    plan.set_timing(0)
    t_iter_ms = loop_ms / (launches * iters)  # one iteration: k_gradient + k_project

    # Breakdown: the same loop without the graph, each kernel bracketed by HIP events (the
    # events and launch gaps make these sum to more than the graph's iteration).
    plan.init()
    prof = plan.profile()
    B = algorithmic_bytes(coefs, W, H)
    achieved = B / (t_iter_ms * 1e-3) / 1e9
    N = W * H
    impl_bytes = sum(28 * N + 12 * c.rounded_px_count for c in coefs)

    plan.init()
    plan.iterate()
    objective, terms = plan.objective()
    out = plan.download()
    finite = all(bool(np.isfinite(p).all()) for p in out)
    plan.close()
    fast = None
    if args.precision == "f32" and not args.no_fast:
        fast = fast_comparison(wl, args, out, objective, B)
    calib = copy_calibration()
    if calib is not None and args.copy_sweep:
        calib["copy_GBps_by_size_MiB"] = {mb: copy_calibration(mb << 20, reps=8)["copy_GBps"] for mb in (16, 64, 128)}
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and args.workload == "tgv_420_2048":
        with open(TRAFFIC_FILE) as fh:
            traffic = json.load(fh).get("bytes_per_iteration_fast" if args.precision == "f32_fast"
                                        else "bytes_per_iteration")

    res = {
        "metric": "Mpixel*iterations/s + time to plain-100-iter objective",
        "value": round(value, 2),
        "unit": "Mpixel*it/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, artefact_amd/synth.py)",
        "config": {
            "workload": describe(wl),
            "width": W, "height": H, "subsampling": subsampling(coefs), "iterations": iters,
            "quality": wl["quality"], "mode": wl["mode"],
            f"time_to_{iters}_iter_objective_ms": round(ms_per_step, 4),
            f"objective_after_{iters}_iter": objective,
            "objective_terms_tv_tgv_prob": [float(t) for t in terms],
            "autotuned": tuning,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "one iteration = k_gradient + k_project (HIP events around the "
                      "timed region's graph launches, divided by the iterations)",
            "iteration_ms": round(t_iter_ms, 5),
            "algorithmic_bytes_per_iteration": B,
            "implementation_bytes_per_iteration": impl_bytes,
            "eager_breakdown_ms": {"k_gradient": round(prof["k_gradient_ms"], 5),
                                   "k_project": round(prof["k_project_ms"], 5)},
            "copy_calibration": calib,
        },
        "finite_output": finite,
    }
    if args.precision != "f32":
        res["dtype"] = f"f32 ({args.precision}: hardware sqrt/reciprocal in the gradient sweep, not bit-exact)"
    if fast is not None:
        res["fast_f32"] = fast
    if extras:
        res["configs"] = other_configs(args, pre)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(wl, args.cpu_iters)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
