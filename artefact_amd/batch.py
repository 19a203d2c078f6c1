"""Many images in one hipGraph: the batched form of the reconstruction loop.

``Batch`` takes one ``compute()`` argument set per image (coefficients, weight,
pweight, iterations, grid, mode), builds every plan, optionally autotunes each
geometry, and captures all images' start images and loops into one graph spread
over several HIP streams.  ``run()`` launches it; ``download(i)`` returns image i's
planes exactly as ``compute()`` would.  No CPU path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _lib
from .compute import _channels, _float_pp, _params
from .types import Coefficient


@dataclass
class BatchImage:
    coefs: Sequence[Coefficient]
    weight: float
    pweight: Sequence[float]
    iterations: int
    max_w: int
    max_h: int
    mode: str = "tgv"
    keep: list = field(default_factory=list, repr=False)


class Batch:
    def __init__(self, images: Sequence[BatchImage], *, streams: int = 0, tune: bool = True):
        """streams = 0: with tune, the fastest of 1-4 streams, timed on the whole graph; else 2."""
        self._lib = _lib.load()
        self.images = list(images)
        arr = (_lib.artefact_batch_image * len(self.images))()
        self._keep: list = []
        for i, im in enumerate(self.images):
            chans, k = _channels(im.coefs)
            p = _params(len(im.coefs), im.weight, im.pweight, im.iterations, im.max_w, im.max_h, im.mode, None,
                        self._keep)
            self._keep += [chans, k, p]
            arr[i].params = ctypes.pointer(p)
            arr[i].channels = ctypes.cast(chans, ctypes.POINTER(_lib.artefact_channel))
        h = ctypes.c_void_p()
        _lib.check(self._lib.artefact_hip_batch_create(arr, len(self.images), streams, 1 if tune else 0,
                                                       ctypes.byref(h)), "artefact_hip_batch_create")
        self._h = h

    @property
    def streams(self) -> int:
        """The number of streams the graph spreads the images over."""
        return int(self._lib.artefact_hip_batch_streams(self._h))

    def run(self) -> None:
        _lib.check(self._lib.artefact_hip_batch_run(self._h), "artefact_hip_batch_run")

    def synchronize(self) -> None:
        _lib.check(self._lib.artefact_hip_batch_synchronize(self._h), "artefact_hip_batch_synchronize")

    def download(self, index: int) -> list[np.ndarray]:
        im = self.images[index]
        out = [np.empty(im.max_w * im.max_h, dtype=np.float32) for _ in im.coefs]
        _lib.check(self._lib.artefact_hip_batch_download(self._h, index, ctypes.cast(_float_pp(out), _lib.FLOAT_PP)),
                   "artefact_hip_batch_download")
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.artefact_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
