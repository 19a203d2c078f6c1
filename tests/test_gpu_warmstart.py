"""Learned warm start with the committed network (artefact_amd/warmstart_weights.npz, trained
on the GPU by `python -m artefact_amd.warmstart` on 20 000 solver-generated 256x256 crops).

What the network gains: on crops outside its training set, K plain iterations from the
CNN's projected prediction end at a lower objective and closer to the converged (300
iteration) result than K plain iterations from the decoded image.  The objective decides:
warm_solve falls back to the plain result when the warm one's objective is higher.
"""
import numpy as np
import pytest

import artefact_amd
from artefact_amd import synth, warmstart

pytestmark = pytest.mark.gpu


def crop(seed, quality, size=256):
    return synth.encode(synth.synth_rgb(size, size, seed), subsampling="420", quality=quality)


def solve(coefs, size, iters, net=None, step_iterations=0):
    with artefact_amd.Plan(coefs, 0.3, [0.001] * 3, iters, size, size) as plan:
        if step_iterations:
            artefact_amd._lib.check(plan._lib.artefact_hip_plan_set_step_iterations(plan._h, step_iterations), "step")
        if net is None:
            plan.init()
        else:
            artefact_amd._lib.check(plan._lib.artefact_hip_plan_warm_init(plan._h, net._h), "warm")
        plan.iterate()
        f, _ = plan.objective()
        return plan.download(), f


@pytest.mark.parametrize("seed,quality", [(900001, 35), (900002, 60), (900003, 80)])
def test_network_beats_the_same_iterations_from_the_decode(seed, quality):
    net = warmstart.WarmStartNet.load()
    coefs = crop(seed, quality)
    print(f"case warm start seed={seed} q={quality}", flush=True)
    ref, _ = solve(coefs, 256, 300)
    cold, f_cold = solve(coefs, 256, 15)
    warm, f_warm = solve(coefs, 256, 15, net)
    net.close()
    d_cold = np.sqrt(np.mean([np.mean((a - r) ** 2) for a, r in zip(cold, ref)]))
    d_warm = np.sqrt(np.mean([np.mean((a - r) ** 2) for a, r in zip(warm, ref)]))
    print(f"objective cold {f_cold:.6g} warm {f_warm:.6g}; rms to 300-iteration result cold {d_cold:.4g} warm {d_warm:.4g}")
    assert f_warm < f_cold
    assert d_warm < d_cold


def test_untrained_network_gains_less():
    """The gain comes from training: an untrained (near-identity) network is worse than the
    trained one after the same iterations."""
    coefs = crop(900004, 50)
    trained = warmstart.WarmStartNet.load()
    untrained = warmstart.WarmStartNet(warmstart.init_params(0))
    _, f_t = solve(coefs, 256, 15, trained)
    _, f_u = solve(coefs, 256, 15, untrained)
    trained.close()
    untrained.close()
    assert f_t < f_u


def test_fallback_returns_the_plain_result():
    coefs = crop(900005, 50, size=128)
    net = warmstart.WarmStartNet.load()
    res = warmstart.warm_solve(coefs, 128, 128, net, iterations=3, plain_iterations=50)
    net.close()
    plain, f_plain = solve(coefs, 128, 50)
    assert res.used_fallback == (res.warm_objective > f_plain)
    if res.used_fallback:
        for a, b in zip(res.planes, plain):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert res.objective <= f_plain


def test_momentum_offset():
    """artefact_hip_plan_set_momentum_offset: offset 0 gives the reference's loop bit for bit (also
    after another offset was set), an offset leaves the first iteration alone (x = x_prev there,
    so the extrapolation factor multiplies zero) and changes the later ones, and offsets above
    2^20 are refused."""
    wl = synth.workload("tgv_420_96x64", seed=4)
    W, H = wl["max_w"], wl["max_h"]
    print("case momentum offset tgv_420_96x64", flush=True)

    def solve(iters, offsets):
        with artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], iters, W, H, mode=wl["mode"]) as p:
            for k in offsets:
                artefact_amd._lib.check(p._lib.artefact_hip_plan_set_momentum_offset(p._h, k), "momentum")
            p.init()
            p.iterate()
            return p.download()

    base = solve(6, [])
    assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(base, solve(6, [40, 0])))
    one = solve(1, []), solve(1, [40])
    assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(*one))
    assert any(not np.array_equal(a, b) for a, b in zip(base, solve(6, [40])))
    with artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], 2, W, H, mode=wl["mode"]) as p:
        assert p._lib.artefact_hip_plan_set_momentum_offset(p._h, (1 << 20) + 1) == -1


def test_step_schedule():
    """artefact_hip_plan_set_step_schedule: first = 0 is set_step_iterations(last) bit for bit, as
    is a schedule from a step to the same step; a decaying schedule changes the result."""
    wl = synth.workload("tgv_420_96x64", seed=5)
    W, H = wl["max_w"], wl["max_h"]
    print("case step schedule tgv_420_96x64", flush=True)

    def solve(setup):
        with artefact_amd.Plan(wl["coefs"], wl["weight"], wl["pweight"], 6, W, H, mode=wl["mode"]) as p:
            setup(p)
            p.init()
            p.iterate()
            return p.download()

    def same(a, b):
        return all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(a, b))

    const = solve(lambda p: artefact_amd._lib.check(p._lib.artefact_hip_plan_set_step_iterations(p._h, 400), "s"))
    assert same(const, solve(lambda p: artefact_amd._lib.check(p._lib.artefact_hip_plan_set_step_schedule(p._h, 0, 400), "s")))
    assert same(const, solve(lambda p: artefact_amd._lib.check(p._lib.artefact_hip_plan_set_step_schedule(p._h, 400, 400), "s")))
    assert not same(const, solve(lambda p: artefact_amd._lib.check(p._lib.artefact_hip_plan_set_step_schedule(p._h, 50, 400), "s")))
