"""CPU reference of the warm-start CNN -- TEST INFRASTRUCTURE ONLY.

The network (no counterpart in the reference): five 3x3 convolutions with zero padding,
channel widths CIN -> 32 -> 32 -> 32 -> 32 -> 1, ReLU after the first four, the output
added to input channel 0 (a residual on the decoded plane).  Activations are NHWC.  This
module restates forward and backward in float64 numpy so that tests can check the HIP
kernels' forward values and parameter gradients; its own gradients are pinned by central
finite differences (tests/test_cnn_oracle.py).  Only tests/ use it.
"""
from __future__ import annotations

import numpy as np

WIDTHS = (2, 32, 32, 32, 32, 1)


def features(x0, quant):
    """Input features of planes x0 (B, H, W) with quantisation tables quant (B, 64), natural order:
    channel 0 = x0 / 64, channel 1 = q[(y mod 8) * 8 + (x mod 8)] / 64 (the table tiled over the
    plane's 8x8 blocks)."""
    x0 = np.asarray(x0, np.float64)
    B, H, W = x0.shape
    q = np.asarray(quant, np.float64).reshape(B, 8, 8)
    qmap = np.tile(q, (1, -(-H // 8), -(-W // 8)))[:, :H, :W]
    return np.stack([x0 / 64.0, qmap / 64.0], axis=-1)


def init_params(seed: int = 0, widths=WIDTHS):
    """He-initialised weights w[l]: (cout, cin, 3, 3) and zero biases b[l]: (cout,)."""
    rng = np.random.default_rng(seed)
    ws, bs = [], []
    for cin, cout in zip(widths[:-1], widths[1:]):
        std = np.sqrt(2.0 / (cin * 9))
        ws.append(rng.normal(0.0, std, size=(cout, cin, 3, 3)))
        bs.append(np.zeros(cout))
    ws[-1] *= 0.1  # start close to the identity map
    return ws, bs


def conv3x3(x, w, b):
    """x: (B, H, W, cin); w: (cout, cin, 3, 3); zero padding; returns (B, H, W, cout)."""
    B, H, W, cin = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((B, H, W, w.shape[0]))
    for dy in range(3):
        for dx in range(3):
            out += xp[:, dy:dy + H, dx:dx + W, :] @ w[:, :, dy, dx].T
    return out + b


def conv3x3_grads(x, w, gout):
    """Gradients of sum(gout * conv3x3(x, w, b)) w.r.t. x, w and b."""
    B, H, W, cin = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    gw = np.zeros_like(w)
    gxp = np.zeros_like(xp)
    for dy in range(3):
        for dx in range(3):
            patch = xp[:, dy:dy + H, dx:dx + W, :].reshape(-1, cin)
            gw[:, :, dy, dx] = gout.reshape(-1, w.shape[0]).T @ patch
            gxp[:, dy:dy + H, dx:dx + W, :] += gout @ w[:, :, dy, dx]
    gb = gout.reshape(-1, w.shape[0]).sum(0)
    return gxp[:, 1:-1, 1:-1, :], gw, gb


def forward(x, ws, bs):
    """Returns (prediction (B, H, W), list of layer inputs, list of pre-activations)."""
    acts, pres = [x], []
    h = x
    for i, (w, b) in enumerate(zip(ws, bs)):
        z = conv3x3(h, w, b)
        pres.append(z)
        h = np.maximum(z, 0.0) if i + 1 < len(ws) else z
        acts.append(h)
    return x[..., 0] + h[..., 0], acts, pres


def loss_and_grads(x, target, ws, bs):
    """Mean squared error over all pixels and its parameter gradients."""
    pred, acts, pres = forward(x, ws, bs)
    n = pred.size
    diff = pred - target
    loss = float((diff ** 2).sum() / n)
    g = (2.0 / n) * diff[..., None]
    gws, gbs = [None] * len(ws), [None] * len(ws)
    for i in range(len(ws) - 1, -1, -1):
        if i + 1 < len(ws):
            g = g * (pres[i] > 0)
        gx, gws[i], gbs[i] = conv3x3_grads(acts[i], ws[i], g)
        g = gx
    return loss, gws, gbs


def adam_step(p, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1), in place on p, m, v."""
    m *= beta1
    m += (1 - beta1) * g
    v *= beta2
    v += (1 - beta2) * g * g
    mhat = m / (1 - beta1 ** t)
    vhat = v / (1 - beta2 ** t)
    p -= lr * mhat / (np.sqrt(vhat) + eps)
