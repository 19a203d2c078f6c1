"""Size-independent checks of loop results, shared by the GPU tests (test infrastructure)."""
from __future__ import annotations

import hashlib

import numpy as np

from artefact_amd import synth


def bits_equal(a, b) -> bool:
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def digest(plane) -> str:
    """SHA-256 of a plane's f32 bytes (little-endian), as tests/golden/make_config_fixtures.py writes it."""
    return hashlib.sha256(np.asarray(plane, np.float32).astype("<f4").tobytes()).hexdigest()


def requant_deviation(planes, coefs, W: int, H: int) -> float:
    """Largest |DCT(down x)/q - c| over every block of every channel: the projection keeps each
    block inside its quantisation interval, so this is <= 1/2 up to f32 rounding.  ``down`` is
    the box mean over the component's sampling factors (compute_projection.rs:14-44)."""
    d = synth.dct_matrix()
    worst = 0.0
    for plane, k in zip(planes, coefs):
        x = np.asarray(plane, np.float64).reshape(H, W)
        hs, vs = k.horizontal_samp_factor, k.vertical_samp_factor
        x = x[: k.rounded_px_h * vs, : k.rounded_px_w * hs]
        if hs > 1 or vs > 1:
            x = x.reshape(k.rounded_px_h, vs, k.rounded_px_w, hs).mean(axis=(1, 3))
        blocks = x.reshape(k.block_h, 8, k.block_w, 8).transpose(0, 2, 1, 3)
        coef = np.einsum("ux,abxy,vy->abuv", d, blocks, d, optimize=True).reshape(-1, 64)
        q = np.asarray(k.quant_table, np.float64)
        dev = np.abs(coef / q - np.asarray(k.dct_coefs, np.float64).reshape(-1, 64))
        worst = max(worst, float(dev.max()))
    return worst


def spread(planes, ref) -> tuple[float, float]:
    """(max-abs difference, share of pixels within 0.05) over all channels, on the 0..255 scale."""
    worst, within, n = 0.0, 0, 0
    for a, b in zip(planes, ref):
        d = np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))
        worst = max(worst, float(d.max()))
        within += int((d <= 0.05).sum())
        n += d.size
    return worst, within / n


def subsample_psnr(planes, sub_file: str, W: int, H: int) -> float:
    """PSNR (peak 255) of planes against a committed subsample of a reference result (an .npz
    of planes p0, p1, ... holding every ``step``-th row and column from ``offset``)."""
    with np.load(sub_file, allow_pickle=False) as z:
        s, o = int(z["step"]), int(z["offset"])
        ref = [z[f"p{c}"].astype(np.float64) for c in range(len(planes))]
    mse = np.mean([np.mean((np.asarray(p, np.float64).reshape(H, W)[o::s, o::s] - r) ** 2)
                   for p, r in zip(planes, ref)])
    return float(10 * np.log10(255.0 ** 2 / mse))
