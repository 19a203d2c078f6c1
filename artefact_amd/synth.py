"""Seeded synthetic inputs for the reconstruction loop.

The reference has no fixtures for this path (its sample JPEGs are generated on
demand, justfile:168-194), so the workloads are defined here:

* ``synth_plane`` -- an 8-bit test image: 128 plus 6 random low-frequency 2D
  cosines (each of amplitude 40/sqrt(6), so the sum has RMS 40/sqrt(2)), 20
  random rectangles with a uniform fill in 0..255, 10 anti-aliased lines of
  width 1..3, Gaussian noise sigma 3, clipped to 0..255 and rounded.
* ``encode`` -- what a baseline JPEG encoder hands the decoder: BT.601 full-range
  YCbCr (the inverse of the reference's output conversion, lib.rs:177-181),
  2x2 box-downsampled chroma for 4:2:0, level shift -128, orthonormal 8x8
  DCT-II per block, and ``round(DCT / Q)`` (half away from zero) as int16, with
  the IJG Annex-K tables scaled to a quality (libjpeg's jpeg_quality_scaling,
  baseline-clamped to 1..255).

Coefficients come out in the layout of backend/artefact-lib/jpeg/mod.rs:8-27:
raster block order, natural order inside a block.
"""
from __future__ import annotations

import numpy as np

from .types import Coefficient

# ITU-T T.81 Annex K, tables K.1 and K.2, natural (row-major) order.
IJG_LUMA = np.array([
    16, 11, 10, 16, 24, 40, 51, 61,
    12, 12, 14, 19, 26, 58, 60, 55,
    14, 13, 16, 24, 40, 57, 69, 56,
    14, 17, 22, 29, 51, 87, 80, 62,
    18, 22, 37, 56, 68, 109, 103, 77,
    24, 35, 55, 64, 81, 104, 113, 92,
    49, 64, 78, 87, 103, 121, 120, 101,
    72, 92, 95, 98, 112, 100, 103, 99,
], dtype=np.int64)
IJG_CHROMA = np.array([
    17, 18, 24, 47, 99, 99, 99, 99,
    18, 21, 26, 66, 99, 99, 99, 99,
    24, 26, 56, 99, 99, 99, 99, 99,
    47, 66, 99, 99, 99, 99, 99, 99,
    99, 99, 99, 99, 99, 99, 99, 99,
    99, 99, 99, 99, 99, 99, 99, 99,
    99, 99, 99, 99, 99, 99, 99, 99,
    99, 99, 99, 99, 99, 99, 99, 99,
], dtype=np.int64)


def quant_table(quality: int, chroma: bool = False) -> np.ndarray:
    """IJG table scaled to ``quality`` (1..100), entries clamped to 1..255."""
    q = int(min(max(quality, 1), 100))
    scale = 5000 // q if q < 50 else 200 - 2 * q
    base = IJG_CHROMA if chroma else IJG_LUMA
    t = (base * scale + 50) // 100
    return np.clip(t, 1, 255).astype(np.uint16)


def dct_matrix() -> np.ndarray:
    """Orthonormal DCT-II matrix D[u, x] = c(u) cos((2x + 1) u pi / 16)."""
    u = np.arange(8)[:, None]
    x = np.arange(8)[None, :]
    d = np.cos((2 * x + 1) * u * np.pi / 16.0)
    d[0, :] *= np.sqrt(1.0 / 8.0)
    d[1:, :] *= np.sqrt(2.0 / 8.0)
    return d


def synth_plane(h: int, w: int, seed: int) -> np.ndarray:
    """8-bit synthetic image (see the module docstring), shape (h, w), uint8."""
    rng = np.random.default_rng(seed)
    yy = np.arange(h, dtype=np.float64)[:, None]
    xx = np.arange(w, dtype=np.float64)[None, :]
    img = np.full((h, w), 128.0)
    amp = 40.0 / np.sqrt(6.0)
    for _ in range(6):
        fx, fy = rng.uniform(-4.0, 4.0, size=2)  # cycles across the image
        ph = rng.uniform(0.0, 2.0 * np.pi)
        img += amp * np.cos(2.0 * np.pi * (fx * xx / w + fy * yy / h) + ph)
    side = max(h, w)
    for _ in range(20):
        rw_ = int(rng.integers(max(side // 32, 1), max(side // 4, 2)))
        rh_ = int(rng.integers(max(side // 32, 1), max(side // 4, 2)))
        x0 = int(rng.integers(0, w))
        y0 = int(rng.integers(0, h))
        img[y0:y0 + rh_, x0:x0 + rw_] = float(rng.integers(0, 256))
    for _ in range(10):
        ax, bx = rng.uniform(0, w, size=2)
        ay, by = rng.uniform(0, h, size=2)
        width = rng.uniform(1.0, 3.0)
        val = float(rng.integers(0, 256))
        pad = width + 1.0
        xa, xb = int(max(min(ax, bx) - pad, 0)), int(min(max(ax, bx) + pad + 1, w))
        ya, yb = int(max(min(ay, by) - pad, 0)), int(min(max(ay, by) + pad + 1, h))
        if xa >= xb or ya >= yb:
            continue
        px = np.arange(xa, xb, dtype=np.float64)[None, :] + 0.5
        py = np.arange(ya, yb, dtype=np.float64)[:, None] + 0.5
        dx, dy = bx - ax, by - ay
        ll = dx * dx + dy * dy
        t = np.clip(((px - ax) * dx + (py - ay) * dy) / ll, 0.0, 1.0) if ll > 0 else np.zeros_like(px + py)
        dist = np.hypot(px - (ax + t * dx), py - (ay + t * dy))
        cov = np.clip(width / 2.0 + 0.5 - dist, 0.0, 1.0)
        sub = img[ya:yb, xa:xb]
        img[ya:yb, xa:xb] = sub * (1.0 - cov) + val * cov
    img += rng.normal(0.0, 3.0, size=(h, w))
    return np.clip(np.round(img), 0, 255).astype(np.uint8)


def rgb_to_ycbcr(r: np.ndarray, g: np.ndarray, b: np.ndarray):
    """BT.601 full range (JFIF), rounded to 8 bits."""
    r, g, b = (a.astype(np.float64) for a in (r, g, b))
    y = 0.299 * r + 0.587 * g + 0.114 * b
    cb = -0.168735892 * r - 0.331264108 * g + 0.5 * b + 128.0
    cr = 0.5 * r - 0.418687589 * g - 0.081312411 * b + 128.0
    return tuple(np.clip(np.round(a), 0, 255).astype(np.uint8) for a in (y, cb, cr))


def downsample2x2(p: np.ndarray) -> np.ndarray:
    """2x2 box average, rounded half up."""
    p = p.astype(np.int64)
    s = p[0::2, 0::2] + p[1::2, 0::2] + p[0::2, 1::2] + p[1::2, 1::2]
    return ((s + 2) // 4).astype(np.uint8)


def downsample(p: np.ndarray, hs: int, vs: int) -> np.ndarray:
    """Box average over hs x vs, rounded half up."""
    if (hs, vs) == (2, 2):
        return downsample2x2(p)
    q = p.astype(np.int64)
    s = sum(q[i::vs, j::hs] for i in range(vs) for j in range(hs))
    n = hs * vs
    return ((s + n // 2) // n).astype(np.uint8)


def _pad_to(p: np.ndarray, h: int, w: int) -> np.ndarray:
    return np.pad(p, ((0, h - p.shape[0]), (0, w - p.shape[1])), mode="edge")


def plane_to_coefficient(plane: np.ndarray, q: np.ndarray, h_samp: int, v_samp: int) -> Coefficient:
    """Level shift, blockwise orthonormal DCT, quantise (half away from zero)."""
    h, w = plane.shape
    assert h % 8 == 0 and w % 8 == 0
    d = dct_matrix()
    blocks = plane.astype(np.float64).reshape(h // 8, 8, w // 8, 8).transpose(0, 2, 1, 3) - 128.0
    coef = np.einsum("ux,abxy,vy->abuv", d, blocks, d, optimize=True)
    qf = q.astype(np.float64).reshape(8, 8)
    v = coef / qf
    qc = np.sign(v) * np.floor(np.abs(v) + 0.5)
    qc = np.clip(qc, -32768, 32767).astype(np.int16)
    return Coefficient(
        rounded_px_w=w, rounded_px_h=h, block_w=w // 8, block_h=h // 8,
        horizontal_samp_factor=h_samp, vertical_samp_factor=v_samp,
        dct_coefs=qc.reshape(-1), quant_table=q.copy(),
    )


def encode(planes, subsampling: str = "420", quality: int = 75) -> list[Coefficient]:
    """Coefficients of an 8-bit image: one plane (grey) or (R, G, B).

    ``subsampling`` is "444", "420", "422" or "440" for colour input.
    """
    if len(planes) == 1:
        p = planes[0]
        h, w = p.shape
        H, W = -(-h // 8) * 8, -(-w // 8) * 8
        return [plane_to_coefficient(_pad_to(p, H, W), quant_table(quality), 1, 1)]
    y, cb, cr = rgb_to_ycbcr(*planes)
    h, w = y.shape
    ql, qc = quant_table(quality), quant_table(quality, chroma=True)
    if subsampling == "444":
        H, W = -(-h // 8) * 8, -(-w // 8) * 8
        return [plane_to_coefficient(_pad_to(a, H, W), t, 1, 1) for a, t in ((y, ql), (cb, qc), (cr, qc))]
    factors = {"420": (2, 2), "422": (2, 1), "440": (1, 2)}
    if subsampling not in factors:
        raise ValueError("subsampling must be '444', '420', '422' or '440'")
    hs, vs = factors[subsampling]
    H, W = -(-h // (8 * vs)) * 8 * vs, -(-w // (8 * hs)) * 8 * hs
    y, cb, cr = (_pad_to(a, H, W) for a in (y, cb, cr))
    return [
        plane_to_coefficient(y, ql, 1, 1),
        plane_to_coefficient(downsample(cb, hs, vs), qc, hs, vs),
        plane_to_coefficient(downsample(cr, hs, vs), qc, hs, vs),
    ]


def synth_rgb(h: int, w: int, seed: int):
    """Three independent seeded planes used as R, G, B."""
    return [synth_plane(h, w, seed * 7919 + 101 * c + 1) for c in range(3)]


def workload(name: str, seed: int = 0) -> dict:
    """The named workloads.

    ``tv_luma_512``   plain TV, luma only, 512x512, quality 50, 50 iterations.
    ``tgv_420_2048``  plain TGV, YCbCr 4:2:0, 2048x1536, quality 75, 100 iterations.
    ``tgv_444_4096``  4:4:4 4096x3072, quality 30, 100 iterations.
    Small variants for tests: ``tgv_<420|422|440|444>_<w>x<h>``, ``tv_luma_<w>x<h>``
    with quality 75 (TGV) / 50 (TV).
    """
    presets = {
        "tv_luma_512": ("tv", 512, 512, "grey", 50, 50),
        "tgv_420_2048": ("tgv", 2048, 1536, "420", 75, 100),
        "tgv_444_4096": ("tgv", 4096, 3072, "444", 30, 100),
    }
    if name in presets:
        mode, w, h, samp, quality, iters = presets[name]
    else:
        kind, sub, dims = name.split("_")
        w, h = (int(v) for v in dims.split("x"))
        mode = kind
        samp = "grey" if sub == "luma" else sub
        quality = 50 if kind == "tv" else 75
        iters = 50 if kind == "tv" else 100
    if samp == "grey":
        coefs = encode([synth_plane(h, w, seed)], quality=quality)
    else:
        coefs = encode(synth_rgb(h, w, seed), subsampling=samp, quality=quality)
    W = max(c.rounded_px_w for c in coefs)
    H = max(c.rounded_px_h for c in coefs)
    return {
        "name": name, "mode": mode, "coefs": coefs, "width": w, "height": h, "max_w": W, "max_h": H,
        "quality": quality, "iterations": iters, "weight": 0.0 if mode == "tv" else 0.3,
        "pweight": [0.001, 0.001, 0.001], "seed": seed,
    }


def batch_specs(count: int = 64, seed: int = 0, min_side: int = 256, max_side: int = 2048,
                iterations: int = 100) -> list[dict]:
    """Parameters of the batched workload's images (see ``batch_workload``), without encoding."""
    rng = np.random.default_rng(seed)
    sides = np.arange(min_side, max_side + 1, 16)
    subs = np.array(["420"] * (count // 2) + ["444"] * (count - count // 2))
    modes = np.array(["tv"] * (count // 2) + ["tgv"] * (count - count // 2))
    rng.shuffle(subs)
    rng.shuffle(modes)
    out = []
    for i in range(count):
        w, h = int(rng.choice(sides)), int(rng.choice(sides))
        quality = int(rng.integers(30, 91))
        out.append({"index": i, "width": w, "height": h, "subsampling": str(subs[i]), "quality": quality,
                    "mode": str(modes[i]), "seed": seed * 1000 + i, "iterations": iterations})
    return out


def batch_image(spec: dict) -> dict:
    """One image of the batched workload from its ``batch_specs`` entry."""
    w, h, mode = spec["width"], spec["height"], spec["mode"]
    coefs = encode(synth_rgb(h, w, spec["seed"]), subsampling=spec["subsampling"], quality=spec["quality"])
    W = max(c.rounded_px_w for c in coefs)
    H = max(c.rounded_px_h for c in coefs)
    return {
        "name": f"batch{spec['index']}_{mode}_{spec['subsampling']}_{w}x{h}_q{spec['quality']}", "mode": mode,
        "coefs": coefs, "width": w, "height": h, "max_w": W, "max_h": H, "quality": spec["quality"],
        "iterations": spec["iterations"], "weight": 0.0 if mode == "tv" else 0.3, "pweight": [0.001, 0.001, 0.001],
        "seed": spec["seed"],
    }


def batch_workload(count: int = 64, seed: int = 0, min_side: int = 256, max_side: int = 2048,
                   iterations: int = 100, workers: int = 1) -> list[dict]:
    """The batched workload: ``count`` images whose width and height are drawn uniformly
    from the multiples of 16 in [min_side, max_side], half 4:2:0 and half 4:4:4 (shuffled),
    IJG quality uniform in 30..90, half plain TV and half TGV (weight 0.3), ``iterations``
    iterations each.  Image i uses generator seed ``seed * 1000 + i``.  With ``workers`` > 1
    the images are synthesised by that many spawned processes (host work only)."""
    specs = batch_specs(count, seed, min_side, max_side, iterations)
    if workers <= 1:
        return [batch_image(s) for s in specs]
    import multiprocessing

    with multiprocessing.get_context("spawn").Pool(workers) as pool:
        return pool.map(batch_image, specs, chunksize=1)
