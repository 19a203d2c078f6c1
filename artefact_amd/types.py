"""Boundary data types of the reconstruction loop.

Mirrors backend/artefact-lib/jpeg/mod.rs:8-27 (``Coefficient``) and the sampling
factor enum of backend/zune-jpeg/src/sample_factor.rs:4-38 (``One``/``Two`` as the
integers 1/2).  The reference stores coefficients and quantisation entries as f32
holding integers (backend/artefact-lib/jpeg/zune.rs:33-45); here they are kept as
the integers themselves (int16 / uint16), which crosses the C ABI bit-exactly.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Coefficient:
    """One colour component: quantised 8x8 DCT blocks plus geometry.

    ``dct_coefs`` holds ``block_count * 64`` values in raster block order
    (block index ``by * block_w + bx``) and natural row-major order inside a
    block (``v * 8 + u``); ``quant_table`` holds 64 entries in natural order.
    ``horizontal_samp_factor`` / ``vertical_samp_factor`` are the upsampling
    factors onto the max grid (1 for luma, 2 for 4:2:0 chroma).
    """

    rounded_px_w: int
    rounded_px_h: int
    block_w: int
    block_h: int
    horizontal_samp_factor: int
    vertical_samp_factor: int
    dct_coefs: np.ndarray
    quant_table: np.ndarray

    def __post_init__(self) -> None:
        self.dct_coefs = _as_int_array(self.dct_coefs, np.int16, "dct_coefs")
        self.quant_table = _as_int_array(self.quant_table, np.uint16, "quant_table")
        if self.quant_table.size != 64:
            raise ValueError("quant_table must hold 64 entries")
        if self.dct_coefs.size != self.block_count * 64:
            raise ValueError(
                f"dct_coefs holds {self.dct_coefs.size} values, expected block_count*64 = {self.block_count * 64}"
            )

    @property
    def rounded_px_count(self) -> int:
        return self.rounded_px_w * self.rounded_px_h

    @property
    def block_count(self) -> int:
        return self.block_w * self.block_h


def _as_int_array(a, dtype, name: str) -> np.ndarray:
    """Converts to ``dtype`` only when every value is represented exactly."""
    arr = np.ascontiguousarray(np.asarray(a).reshape(-1))
    if arr.dtype == dtype:
        return arr
    info = np.iinfo(dtype)
    as_f = arr.astype(np.float64)
    if not np.all(np.isfinite(as_f)) or np.any(as_f != np.round(as_f)) or as_f.min(initial=0) < info.min or as_f.max(initial=0) > info.max:
        raise ValueError(f"{name} holds values that are not exact {np.dtype(dtype).name} integers")
    return np.ascontiguousarray(arr.astype(dtype))


def max_grid(coefs) -> tuple[int, int, int]:
    """(max_rounded_px_w, max_rounded_px_h, count), as backend/artefact-lib/lib.rs:108-116."""
    w = max(c.rounded_px_w for c in coefs)
    h = max(c.rounded_px_h for c in coefs)
    return w, h, w * h
