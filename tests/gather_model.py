"""numpy model of the GPU kernels' gather-form gradient (CPU only).

The HIP kernel k_gradient evaluates, at every destination pixel q, the terms the
reference scatters into q (compute_step_tv.rs:74-86, compute_step_tv2.rs:96-135)
in the reference's accumulation order:
  prob(q); TV from q-W, q-1, q; second order from q-W, q-W+1, q-1, q, q+1, q+W-1, q+W
with zero-valued terms for sources outside the image or with a zero norm.  This
module states the same thing with whole-plane float32 numpy operations, so the
tests can check the gather form against the oracle's scatter form bit for bit.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def _shift(a, dy, dx):
    """out[y, x] = a[y + dy, x + dx], 0 where that is outside the plane."""
    H, W = a.shape
    out = np.zeros_like(a)
    ys = slice(max(0, -dy), min(H, H - dy))
    xs = slice(max(0, -dx), min(W, W - dx))
    yd = slice(max(0, dy), min(H, H + dy))
    xd = slice(max(0, dx), min(W, W + dx))
    out[ys, xs] = a[yd, xd]
    return out


def forward_diffs(y):
    H, W = y.shape
    gx = np.zeros_like(y)
    gy = np.zeros_like(y)
    gx[:, :-1] = y[:, 1:] - y[:, :-1]
    gy[:-1, :] = y[1:, :] - y[:-1, :]
    return gx, gy


def gradient(planes, alpha_tv, alpha_tgv, prob=None):
    """prob + TV + second-order gradient, gather form; planes: list of (H, W) f32."""
    C = len(planes)
    diffs = [forward_diffs(p) for p in planes]
    H, W = planes[0].shape
    n = np.zeros((H, W), f32)
    for gx, gy in diffs:
        n = n + gx * gx
        n = n + gy * gy
    n = np.sqrt(n)
    terms = []
    n2 = np.zeros((H, W), f32)
    for gx, gy in diffs:
        gxx = np.zeros_like(gx)
        gyx = np.zeros_like(gx)
        gyy = np.zeros_like(gx)
        gxy = np.zeros_like(gx)
        gxx[:, 1:] = gx[:, 1:] - gx[:, :-1]
        gyx[:, 1:] = gy[:, 1:] - gy[:, :-1]
        gyy[1:, :] = gy[1:, :] - gy[:-1, :]
        gxy[1:, :] = gx[1:, :] - gx[:-1, :]
        sym = ((gxy.astype(np.float64) + gyx.astype(np.float64)) * 0.5).astype(f32)
        n2 = n2 + (gyy * gyy + (f32(2) * (sym * sym) + gxx * gxx))
        terms.append((gxx, gyy, sym))
    n2 = np.sqrt(n2)
    a1, a2 = f32(alpha_tv), f32(alpha_tgv)
    nz = n != 0
    nz2 = n2 != 0
    safe_n = np.where(nz, n, f32(1))
    safe_n2 = np.where(nz2, n2, f32(1))
    out = []
    for c in range(C):
        gx, gy = diffs[c]
        gxx, gyy, s = terms[c]
        tvs = np.where(nz, (a1 * -(gx + gy)) / safe_n, f32(0))
        tvr = np.where(nz, (a1 * gx) / safe_n, f32(0))
        tvd = np.where(nz, (a1 * gy) / safe_n, f32(0))
        tgs = np.where(nz2, a2 * (-(f32(2) * gyy + (f32(2) * gxx + f32(2) * s)) / safe_n2), f32(0))
        tgx = np.where(nz2, a2 * ((s + gxx) / safe_n2), f32(0))
        tgy = np.where(nz2, a2 * ((gyy + s) / safe_n2), f32(0))
        tgd = np.where(nz2, a2 * ((-s) / safe_n2), f32(0))
        g = np.zeros((H, W), f32)
        if prob is not None:
            g = g + prob[c]
        g = g + _shift(tvd, -1, 0)
        g = g + _shift(tvr, 0, -1)
        g = g + tvs
        g = g + _shift(tgy, -1, 0)
        g = g + _shift(tgd, -1, 1)
        g = g + _shift(tgx, 0, -1)
        g = g + tgs
        g = g + _shift(tgx, 0, 1)
        g = g + _shift(tgd, 1, -1)
        g = g + _shift(tgy, 1, 0)
        out.append(g)
    return out
