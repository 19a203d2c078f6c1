"""The GPU kernels' gather form of the gradient equals the reference's scatter
form bit for bit (CPU: numpy model of k_gradient vs the oracle)."""
import numpy as np
import pytest

import oracle
from artefact_amd import synth
from gather_model import gradient as gather_gradient

f32 = np.float32


def prob_planes(coefs, cos, pweight, W, H):
    """alpha_p * IDCT((cos - c q) / q^2), replicated onto the grid (compute_step_prob.rs:14-58)."""
    out = []
    for k, cs, pw in zip(coefs, cos, pweight):
        g = np.zeros((H, W), f32)
        if pw != 0:
            q = np.tile(k.quant_table.astype(f32), k.block_count)
            d = (cs - k.dct_coefs.astype(f32) * q) / (q * q)
            pix = oracle.idct8x8(d.reshape(-1, 64))
            ap = f32(pw) * f32(2) * f32(255) * np.sqrt(f32(2))
            img = oracle.unboxing(ap * pix, k.rounded_px_w, k.block_w, k.block_h).reshape(k.rounded_px_h, k.rounded_px_w)
            up = np.repeat(np.repeat(img, k.vertical_samp_factor, 0), k.horizontal_samp_factor, 1)
            g[: up.shape[0], : up.shape[1]] = up
        out.append(g)
    return out


@pytest.mark.parametrize("name,seed", [("tgv_420_96x64", 1), ("tgv_444_40x24", 2), ("tv_luma_48x32", 3), ("tgv_420_16x16", 4)])
def test_gather_equals_scatter(name, seed):
    wl = synth.workload(name, seed=seed)
    W, H = wl["max_w"], wl["max_h"]
    C = len(wl["coefs"])
    rng = np.random.default_rng(seed)
    planes = [(oracle.init_plane(k, W, H) + rng.normal(size=W * H).astype(f32) * f32(4)).astype(f32) for k in wl["coefs"]]
    cos = [(k.dct_coefs.astype(f32) * np.tile(k.quant_table.astype(f32), k.block_count)
            + rng.normal(size=k.rounded_px_count).astype(f32)).astype(f32) for k in wl["coefs"]]
    weight = wl["weight"]
    pw = [0.001, 0.002, 0.003][:C]
    g, _, _ = oracle.gradient(wl["coefs"], weight, pw, W, H, planes, cos)
    atv = f32(1) / np.sqrt(f32(C))
    atg = (f32(weight) / np.sqrt(f32(2))) / np.sqrt(f32(C))
    gm = gather_gradient([p.reshape(H, W) for p in planes], atv, atg, prob_planes(wl["coefs"], cos, pw, W, H))
    for a, b in zip(g, gm):
        assert np.array_equal(a.view(np.uint32), b.reshape(-1).view(np.uint32))


def _widened_midpoint(a, b):
    """f32::midpoint as the reference restates it: widened to f64, halved, rounded once."""
    return ((a.astype(np.float64) + b.astype(np.float64)) * 0.5).astype(f32)


def test_f32_midpoint_equals_widened_form():
    """k_gradient forms the symmetric second difference as fl32(fl32(a + b) * 0.5); that equals
    the widened f64 form for every pair whose sum stays below 2^128.  Checked on random bit
    patterns over the whole exponent range (subnormals included), on cancelling pairs, on tiny
    pairs whose half is subnormal and on pairs with wide exponent gaps."""
    rng = np.random.default_rng(7)
    n = 1 << 21
    bits = rng.integers(0, 1 << 32, size=(2, n), dtype=np.uint64).astype(np.uint32)
    a, b = bits.view(f32)
    keep = np.isfinite(a) & np.isfinite(b) & (np.abs(a) < f32(1.5e38)) & (np.abs(b) < f32(1.5e38))
    cases = [(a[keep], b[keep])]
    # cancelling pairs: b = -a nudged by a few ulps
    x = rng.uniform(-1e30, 1e30, n).astype(f32)
    nudge = rng.integers(-4, 5, n).astype(np.int32)
    y = (-x).view(np.int32) + nudge
    cases.append((x, y.view(f32)))
    # tiny operands: sums and halves in and around the subnormal range
    t = (rng.integers(-(1 << 26), 1 << 26, size=(2, n)).astype(np.float64) * 2.0 ** -149).astype(f32)
    cases.append((t[0], t[1]))
    # wide exponent gaps
    big = rng.uniform(-1e6, 1e6, n).astype(f32)
    small = (rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-45, -2, n)).astype(f32)
    cases.append((big, small))
    for p, q in cases:
        got = (p + q) * f32(0.5)
        want = _widened_midpoint(p, q)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
