"""GPU parity: the HIP path through the C ABI against the CPU oracle.

The HIP kernels reproduce the reference's f32 operation order; the only order
difference is the gradient norm (sequential f32 fold in the reference,
compute_step.rs:54-58; on the GPU, per-lane f64 sums merged per workgroup in double-double
and added exactly in a fixed-point accumulator, rounded once).  So:

* against the oracle in norm_mode="f64" (same norm) the planes must be
  bit-identical (``array_equal`` on the raw bits);
* against the oracle in the reference's own norm mode they must agree within
  the tolerances stated below (max-abs on the 0..255 scale, relative objective).
"""
import math

import numpy as np
import pytest

import artefact_amd
import oracle
from artefact_amd import synth

pytestmark = pytest.mark.gpu

# Workloads: (name, iterations).  Widths that are not multiples of the 64-pixel
# tile exercise partial tiles; 4:4:4, 4:2:0 and luma-only cover every sampling path.
CASES = [
    ("tgv_420_96x64", 0),
    ("tgv_420_96x64", 1),
    ("tgv_420_96x64", 2),
    ("tgv_420_96x64", 12),
    ("tgv_444_72x40", 7),
    ("tv_luma_64x64", 9),
    ("tv_luma_8x8", 3),
    ("tgv_420_16x16", 5),
    ("tgv_420_208x144", 20),
]


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def run_gpu(wl, iters, **kw):
    W, H = wl["max_w"], wl["max_h"]
    C = len(wl["coefs"])
    return artefact_amd.compute(C, wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, W * H,
                                mode=wl["mode"], **kw)


@pytest.mark.parametrize("name,iters", CASES)
def test_bit_exact_vs_oracle_f64_norm(name, iters):
    wl = synth.workload(name, seed=11)
    W, H = wl["max_w"], wl["max_h"]
    print(f"case {name} iters={iters}", flush=True)
    gpu = run_gpu(wl, iters)
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, norm_mode="f64")
    for c, (a, b) in enumerate(zip(gpu, ref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"


def test_first_gradient_bit_exact():
    wl = synth.workload("tgv_420_96x64", seed=5)
    W, H = wl["max_w"], wl["max_h"]
    with artefact_amd.Plan(wl["coefs"], 0.3, wl["pweight"], 10, W, H) as plan:
        plan.init()
        g = plan.first_gradient()
        x0 = plan.download()
    init = [oracle.init_plane(c, W, H) for c in wl["coefs"]]
    for a, b in zip(x0, init):
        assert bits_equal(a, b)
    cos = [c.dct_coefs.astype(np.float32) * np.tile(c.quant_table.astype(np.float32), c.block_count) for c in wl["coefs"]]
    gref, _, _ = oracle.gradient(wl["coefs"], 0.3, wl["pweight"], W, H, init, cos)
    for c, (a, b) in enumerate(zip(g, gref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"


def test_norms_and_cos_match_oracle():
    wl = synth.workload("tgv_420_96x64", seed=2)
    W, H = wl["max_w"], wl["max_h"]
    iters = 6
    with artefact_amd.Plan(wl["coefs"], 0.3, wl["pweight"], iters, W, H) as plan:
        plan.init()
        plan.iterate()
        out = plan.download()
        norms = plan.last_norms()
        cos = plan.download_cos()
    ref, rnorms = oracle.compute(wl["coefs"], 0.3, wl["pweight"], iters, W, H, norm_mode="f64", return_norms=True)
    assert bits_equal(norms, rnorms[-1])
    for a, b in zip(out, ref):
        assert bits_equal(a, b)
    # cos holds the clamped coefficients of the last projection: inside the box.
    for c, k in zip(cos, wl["coefs"]):
        q = np.tile(k.quant_table.astype(np.float32), k.block_count)
        cf = k.dct_coefs.astype(np.float32)
        assert np.all(c >= (cf - 0.5) * q) and np.all(c <= (cf + 0.5) * q)


def test_tolerance_vs_reference_norm():
    """Against the reference's own sequential-f32 norm: the loop is not bit-exact
    there (the reference's norm itself carries ~1e-6..1e-4 relative rounding), so the
    bar is max-abs <= 0.05 on the 0..255 scale and relative objective error <= 1e-4."""
    wl = synth.workload("tgv_420_208x144", seed=4)
    W, H = wl["max_w"], wl["max_h"]
    iters = 30
    gpu = run_gpu(wl, iters)
    ref = oracle.compute(wl["coefs"], 0.3, wl["pweight"], iters, W, H, norm_mode="reference")
    worst = max(float(np.abs(a - b).max()) for a, b in zip(gpu, ref))
    fg, _ = oracle.objective(wl["coefs"], 0.3, wl["pweight"], W, H, gpu)
    fr, _ = oracle.objective(wl["coefs"], 0.3, wl["pweight"], W, H, ref)
    assert worst <= 0.05, worst
    assert abs(fg - fr) / abs(fr) <= 1e-4


def test_plan_graph_replay_matches_eager():
    wl = synth.workload("tgv_420_96x64", seed=8)
    W, H = wl["max_w"], wl["max_h"]
    outs = []
    for graph in (False, True):
        with artefact_amd.Plan(wl["coefs"], 0.3, wl["pweight"], 9, W, H, graph=graph) as plan:
            for _ in range(3):
                plan.init()
                plan.iterate()
                outs.append(plan.download())
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert bits_equal(a, b)


def test_requantises_at_full_size():
    """Every output block re-quantises to its input coefficients (size-independent
    property at the 2048x1536 4:2:0 workload): |DCT(down x)/q - c| <= 1/2 + eps."""
    wl = synth.workload("tgv_420_2048", seed=0)
    W, H = wl["max_w"], wl["max_h"]
    out = run_gpu(wl, 10)
    d = synth.dct_matrix()
    for plane, k in zip(out, wl["coefs"]):
        x = plane.reshape(H, W).astype(np.float64)
        if k.horizontal_samp_factor == 2:
            x = 0.25 * (x[0::2, 0::2] + x[1::2, 0::2] + x[0::2, 1::2] + x[1::2, 1::2])
        blocks = x.reshape(k.block_h, 8, k.block_w, 8).transpose(0, 2, 1, 3)
        coef = np.einsum("ux,abxy,vy->abuv", d, blocks, d, optimize=True).reshape(-1, 64)
        q = k.quant_table.astype(np.float64)
        dev = np.abs(coef / q - k.dct_coefs.reshape(-1, 64))
        assert dev.max() <= 0.5 + 1e-3, dev.max()


@pytest.mark.parametrize("name,iters", [("tgv_422_80x40", 9), ("tgv_440_48x64", 9), ("tgv_420_208x144", 7)])
def test_generic_projection_path_bit_exact(name, iters, monkeypatch):
    """Channel layouts without a fused projection kernel (4:2:2, 4:4:0) and the
    per-channel kernel forced on 4:2:0."""
    monkeypatch.setenv("ARTEFACT_GENERIC_PROJECT", "1" if name.startswith("tgv_420") else "0")
    wl = synth.workload(name, seed=13)
    W, H = wl["max_w"], wl["max_h"]
    print(f"case {name} iters={iters}", flush=True)
    gpu = run_gpu(wl, iters)
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, norm_mode="f64")
    for c, (a, b) in enumerate(zip(gpu, ref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"


def test_component_smaller_than_grid_bit_exact():
    """A component whose plane is smaller than the grid without subsampling
    (resampled with 1x1 boxes, compute_projection.rs:14) and a grid larger than
    every component (get_fdata's clamp, coef.rs:157-160)."""
    from artefact_amd.types import Coefficient
    wl = synth.workload("tgv_444_48x32", seed=21)
    y, cb, cr = wl["coefs"]
    small = Coefficient(32, 24, 4, 3, 1, 1, cb.dct_coefs.reshape(4, 6, 64)[:3, :4].reshape(-1), cb.quant_table)
    coefs = [y, small, cr]
    W, H = 56, 40
    gpu = artefact_amd.compute(3, coefs, 0.3, [0.001, 0.002, 0.003], 8, W, H, W * H)
    ref = oracle.compute(coefs, 0.3, [0.001, 0.002, 0.003], 8, W, H, norm_mode="f64")
    for c, (a, b) in enumerate(zip(gpu, ref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"


@pytest.mark.parametrize("name,iters", [("tgv_420_96x64", 6), ("tgv_444_72x40", 4), ("tv_luma_64x64", 5), ("tgv_422_80x40", 3)])
def test_objective_matches_oracle(name, iters):
    """Row O: the GPU f64 objective of the returned iterate equals the oracle's to 1e-12."""
    wl = synth.workload(name, seed=17)
    W, H = wl["max_w"], wl["max_h"]
    with artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, mode=wl["mode"]) as plan:
        plan.init()
        f0, t0 = plan.objective()
        plan.iterate()
        out = plan.download()
        f1, t1 = plan.objective()
    init = [oracle.init_plane(k, W, H) for k in wl["coefs"]]
    r0, rt0 = oracle.objective(wl["coefs"], wl["weight"], wl["pweight"], W, H, init)
    r1, rt1 = oracle.objective(wl["coefs"], wl["weight"], wl["pweight"], W, H, out)
    np.testing.assert_allclose(t0, rt0, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(t1, rt1, rtol=1e-12, atol=1e-9)
    assert f1 < f0  # the loop decreases the objective from the decoded start


@pytest.mark.parametrize("rows", [2, 3, 4, 5, 7, 300])
@pytest.mark.parametrize("name", ["tgv_420_208x144", "tv_luma_64x64"])
def test_strip_heights_bit_exact(name, rows, monkeypatch):
    """k_gradient's row loop leaves its three-step body after step 1, 2 or 3 depending on
    the strip height mod 3; a strip taller than the image leaves most waves empty."""
    monkeypatch.setenv("ARTEFACT_KG_ROWS", str(rows))
    wl = synth.workload(name, seed=rows)
    W, H = wl["max_w"], wl["max_h"]
    print(f"case {name} rows={rows}", flush=True)
    gpu = run_gpu(wl, 4)
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], 4, W, H, norm_mode="f64")
    for c, (a, b) in enumerate(zip(gpu, ref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"


def same_bits_or_both_nan(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


@pytest.mark.parametrize("scale", [1e-38, 1e17])
def test_first_gradient_extreme_scales(scale):
    """Start images whose differences are subnormal (the midpoint's widened path) or huge
    (squared norms near 1e36) give the oracle's gradient bit for bit.  Squared differences
    that overflow f32 are outside the contract: int16 coefficients and u16 tables cannot
    produce them, and the kernel does not reproduce a / inf = 0."""
    wl = synth.workload("tgv_444_64x40", seed=3)
    W, H = wl["max_w"], wl["max_h"]
    rng = np.random.default_rng(5)
    planes = []
    for _ in range(3):
        p = (rng.standard_normal((H, W)) * scale).astype(np.float32)
        p[:, ::7] = 0.0
        planes.append(p.reshape(-1))
    print(f"case extreme scale={scale}", flush=True)
    with artefact_amd.Plan(wl["coefs"], 0.3, wl["pweight"], 1, W, H, init_override=planes) as plan:
        plan.init()
        g = plan.first_gradient()
        norms = plan.last_norms()
    cos = [c.dct_coefs.astype(np.float32) * np.tile(c.quant_table.astype(np.float32), c.block_count) for c in wl["coefs"]]
    gref, _, _ = oracle.gradient(wl["coefs"], 0.3, wl["pweight"], W, H, planes, cos)
    for c, (a, b) in enumerate(zip(g, gref)):
        assert same_bits_or_both_nan(a, b), f"channel {c}"
    # The norms from the fixed-point accumulator: the square root of the correctly rounded
    # sum of squares (math.fsum), rounded to f32.
    for c, b in enumerate(gref):
        exact = math.fsum((np.asarray(b, np.float64) ** 2).tolist())
        assert bits_equal(np.float32(np.sqrt(exact)), norms[c]), f"channel {c}: {norms[c]!r} vs {np.sqrt(exact)!r}"


@pytest.mark.parametrize("kernel", ["lds", "lane"])
@pytest.mark.parametrize("name,iters", [("tgv_420_208x144", 6), ("tgv_444_72x40", 5), ("tv_luma_64x64", 4),
                                        ("tgv_420_96x64", 1)])
def test_both_fused_projection_kernels_bit_exact(kernel, name, iters, monkeypatch):
    """The fused layouts have two projection kernels (LDS transposes / lane transposes); both
    give the oracle's bits."""
    monkeypatch.setenv("ARTEFACT_PROJECT_KERNEL", kernel)
    wl = synth.workload(name, seed=23)
    W, H = wl["max_w"], wl["max_h"]
    print(f"case {name} iters={iters} kernel={kernel}", flush=True)
    gpu = run_gpu(wl, iters)
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, norm_mode="f64")
    for c, (a, b) in enumerate(zip(gpu, ref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"


def test_timed_launches_same_bits_and_counted():
    """Event pairs around iterate() (set_timing) change nothing in the result; the pairs are
    counted, and launches past the requested number are not timed."""
    wl = synth.workload("tgv_420_96x64", seed=29)
    W, H = wl["max_w"], wl["max_h"]
    print("case timing tgv_420_96x64 iters=6", flush=True)
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], 6, W, H, norm_mode="f64")
    with artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], 6, W, H, mode=wl["mode"], graph=True) as plan:
        plan.set_timing(2)
        for _ in range(3):
            plan.init()
            plan.iterate()
            out = plan.download()
            for c, (a, b) in enumerate(zip(out, ref)):
                assert bits_equal(a, b), f"channel {c}"
        ms, n = plan.timing()
        assert n == 2
        assert ms > 0.0
        plan.set_timing(0)
        assert plan.timing() == (0.0, 0)


@pytest.mark.parametrize("name,iters", [("tgv_420_2048", 3), ("tgv_444_4096", 2), ("tv_luma_512", 6)])
def test_bit_exact_at_configured_sizes(name, iters):
    """The configured workloads at their real shapes (default strip height and projection
    kernel for each), a few iterations: bit-identical to the oracle with the f64 norm."""
    wl = synth.workload(name, seed=0)
    W, H = wl["max_w"], wl["max_h"]
    print(f"case {name} {W}x{H} iters={iters}", flush=True)
    gpu = run_gpu(wl, iters)
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, norm_mode="f64")
    for c, (a, b) in enumerate(zip(gpu, ref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"


@pytest.mark.parametrize("name,iters,env", [
    ("tgv_420_208x144", 7, {}),
    ("tgv_420_96x64", 5, {"ARTEFACT_PROJECT_KERNEL": "lds"}),
    ("tgv_444_72x40", 6, {}),
    ("tv_luma_64x64", 6, {"ARTEFACT_GENERIC_PROJECT": "1"}),
    ("tgv_422_80x40", 5, {}),
    ("tgv_420_2048", 2, {}),
])
def test_projection_kernels_bit_exact(name, iters, env, monkeypatch):
    """Each projection kernel forms the norms from the accumulator in every workgroup: the
    same bits as the oracle with the f64 norm."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    wl = synth.workload(name, seed=31)
    W, H = wl["max_w"], wl["max_h"]
    print(f"case projection kernel {name} iters={iters} {env}", flush=True)
    gpu = run_gpu(wl, iters)
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, norm_mode="f64")
    for c, (a, b) in enumerate(zip(gpu, ref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"


@pytest.mark.parametrize("name", ["tgv_444_72x40", "tgv_420_96x64"])
def test_two_channel_plan_bit_exact(name):
    """Two channels (the luma and the first chroma component of a colour workload): the C = 2
    instances of the gradient sweep, the norm accumulator and the per-channel projection."""
    wl = synth.workload(name, seed=43)
    W, H = wl["max_w"], wl["max_h"]
    coefs = wl["coefs"][:2]
    print(f"case two channels {name}", flush=True)
    gpu = artefact_amd.compute(2, coefs, wl["weight"], wl["pweight"], 6, W, H, W * H, mode=wl["mode"])
    ref = oracle.compute(coefs, wl["weight"], wl["pweight"], 6, W, H, norm_mode="f64")
    assert len(gpu) == 2
    for c, (a, b) in enumerate(zip(gpu, ref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"


@pytest.mark.parametrize("wgs", ["0", "1", "7", "40"])
@pytest.mark.parametrize("name", ["tgv_420_208x144", "tgv_444_72x40", "tv_luma_64x64"])
def test_projection_workgroup_counts_bit_exact(name, wgs, monkeypatch):
    """The lane projection kernel with one unit per wave (0) and with 1, 7 or 40 workgroups,
    whose waves then loop over several units each: the oracle's bits."""
    monkeypatch.setenv("ARTEFACT_LANE_WGS", wgs)
    monkeypatch.setenv("ARTEFACT_PROJECT_KERNEL", "lane")
    wl = synth.workload(name, seed=41)
    W, H = wl["max_w"], wl["max_h"]
    print(f"case projection workgroups {name} wgs={wgs}", flush=True)
    gpu = run_gpu(wl, 4)
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], 4, W, H, norm_mode="f64")
    for c, (a, b) in enumerate(zip(gpu, ref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"


@pytest.mark.parametrize("precision", ["f32", "bf16x3"])
def test_norms_independent_of_strip_height(precision, monkeypatch):
    """Strips of 2 rows and the default height give k_gradient different workgroup counts and
    finishing orders; the accumulated norms and the planes keep their bits."""
    wl = synth.workload("tgv_444_72x40", seed=37)
    W, H = wl["max_w"], wl["max_h"]
    outs, norms = [], []
    for rows in ("2", None):
        if rows:
            monkeypatch.setenv("ARTEFACT_KG_ROWS", rows)
        else:
            monkeypatch.delenv("ARTEFACT_KG_ROWS", raising=False)
        print(f"case norms {precision} rows={rows}", flush=True)
        with artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], 6, W, H, mode=wl["mode"]) as plan:
            plan.set_precision(precision)
            plan.init()
            plan.iterate()
            outs.append(plan.download())
            norms.append(plan.last_norms())
    for a, b in zip(outs[0], outs[1]):
        assert bits_equal(a, b)
    assert bits_equal(norms[0], norms[1])


def test_golden_fixtures_bit_exact():
    """The committed golden fixtures (tests/golden: 4:2:0, 4:4:4, 4:2:2, 4:4:0 and luma-only
    TV, partial tiles) through the C ABI: the stored bits of the oracle's planes."""
    from golden_cases import loop_cases

    for case in loop_cases():
        print(f"case golden {case['name']} iters={case['iterations']}", flush=True)
        W, H = case["max_w"], case["max_h"]
        gpu = artefact_amd.compute(len(case["coefs"]), case["coefs"], case["weight"], case["pweight"],
                                   case["iterations"], W, H, W * H, mode=case["mode"])
        for c, (a, b) in enumerate(zip(gpu, case["planes"])):
            assert bits_equal(a, b), f"{case['name']} channel {c}: max abs diff {np.abs(a - b).max()}"


def test_concurrent_compute_calls_match_sequential():
    """Separate mode in the reference calls compute(1, ...) for each channel concurrently from
    rayon workers (lib.rs:140-149).  Four host threads, each with its own one-channel image,
    call the C ABI at once; every result must equal the oracle's bits, as sequential calls
    do."""
    from concurrent.futures import ThreadPoolExecutor

    jobs = []
    for seed in range(4):
        wl = synth.workload("tgv_420_96x64", seed=30 + seed)
        k = wl["coefs"][seed % 3]  # luma or an upsampled chroma component, on the image's grid
        jobs.append((k, wl["max_w"], wl["max_h"]))

    def run(job):
        k, W, H = job
        return artefact_amd.compute(1, [k], 0.3, [0.001] * 3, 12, W, H, W * H)

    print("case concurrent compute x4", flush=True)
    with ThreadPoolExecutor(max_workers=4) as pool:
        results = list(pool.map(run, jobs * 2))
    for (k, W, H), got in zip(jobs * 2, results):
        ref = oracle.compute([k], 0.3, [0.001] * 3, 12, W, H, norm_mode="f64")
        assert bits_equal(got[0], ref[0])


@pytest.mark.parametrize("name,iters", [("tgv_420_96x64", 6), ("tv_luma_64x64", 5), ("tgv_444_72x40", 4)])
def test_reference_default_init_through_override(name, iters):
    """The reference's default build runs pipeline_scalar, whose start image unboxes inside the
    block loop (pipeline_scalar/coef.rs:29-50).  A caller that wants that start image passes it
    as init_override; the loop from it equals the oracle's loop from the same quirk init."""
    wl = synth.workload(name, seed=17)
    W, H = wl["max_w"], wl["max_h"]
    print(f"case scalar-quirk init {name} iters={iters}", flush=True)
    init = [oracle.init_plane(k, W, H, init_mode="scalar_quirk") for k in wl["coefs"]]
    gpu = run_gpu(wl, iters, init_override=init)
    ref = oracle.compute(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, init_mode="scalar_quirk",
                         norm_mode="f64")
    for c, (a, b) in enumerate(zip(gpu, ref)):
        assert bits_equal(a, b), f"channel {c}: max abs diff {np.abs(a - b).max()}"
