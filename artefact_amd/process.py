"""Host-side mirror of ``Artefact`` (backend/artefact-lib/lib.rs:34-212).

The builder keeps the reference's names and defaults (lib.rs:49-70): ``weight``,
``pweight`` and ``iterations`` take a ``ForAll`` / ``ForEach`` value collection,
``separate_components`` and ``benchmark`` are flags, ``source`` names the input.
``process()`` returns the RGB image as a uint8 array of shape (height, width, 3)
(the reference returns an ``image::RgbImage``), runs the loop and the output stage
on the GPU through ``artefact_hip_process``, and raises where the reference returns
``Err``: ``ArtefactProcessError("Source is not set")``, ``"Failed to read JPEG: …"``
and ``"BENCHMARK"`` (lib.rs:100-157).
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Generic, Sequence, TypeVar, Union

import numpy as np

from . import _lib
from .compute import _channels
from .types import Coefficient

T = TypeVar("T")


@dataclasses.dataclass(frozen=True)
class ForAll(Generic[T]):
    """``ValueCollection::ForAll``: one value for every channel (lib.rs:34-38)."""
    value: T

    def to_slice(self) -> list:
        return [self.value] * 3


@dataclasses.dataclass(frozen=True)
class ForEach(Generic[T]):
    """``ValueCollection::ForEach``: one value per channel (lib.rs:34-38)."""
    values: tuple

    def __post_init__(self):
        if len(self.values) != 3:
            raise ValueError("ForEach takes exactly three values")

    def to_slice(self) -> list:
        return list(self.values)


ValueCollection = Union[ForAll, ForEach]


@dataclasses.dataclass(frozen=True)
class JpegSource:
    """``JpegSource`` (jpeg/mod.rs:39-43): a file path or the bytes of a JPEG."""
    path: str | None = None
    buffer: bytes | None = None

    @staticmethod
    def File(path: str) -> "JpegSource":
        return JpegSource(path=path)

    @staticmethod
    def Buffer(data: bytes) -> "JpegSource":
        return JpegSource(buffer=bytes(data))


@dataclasses.dataclass
class Jpeg:
    """``Jpeg`` (jpeg/mod.rs:29-36): the decoded coefficient set of one image."""
    nchannel: int
    real_px_w: int
    real_px_h: int
    coefs: list[Coefficient]


class ArtefactProcessError(RuntimeError):
    """The ``Err(String)`` of ``Artefact::process``."""


class Artefact:
    """Builder of lib.rs:49-97 with the defaults of lib.rs:60-71."""

    def __init__(self):
        self._weight: ValueCollection = ForAll(0.3)
        self._pweight: ValueCollection = ForAll(0.001)
        self._iterations: ValueCollection = ForAll(50)
        self._separate_components = False
        self._benchmark = False
        self._source: Union[JpegSource, Jpeg, None] = None

    def source(self, source: Union[JpegSource, Jpeg]) -> "Artefact":
        """A ``JpegSource``, or an already decoded ``Jpeg`` (skips the decoder)."""
        self._source = source
        return self

    def weight(self, weight: ValueCollection) -> "Artefact":
        self._weight = weight
        return self

    def pweight(self, pweight: ValueCollection) -> "Artefact":
        self._pweight = pweight
        return self

    def iterations(self, iterations: ValueCollection) -> "Artefact":
        self._iterations = iterations
        return self

    def benchmark(self, benchmark: bool) -> "Artefact":
        self._benchmark = bool(benchmark)
        return self

    def separate_components(self, separate_components: bool) -> "Artefact":
        self._separate_components = bool(separate_components)
        return self

    def _decoded(self) -> Jpeg:
        if self._source is None:
            raise ArtefactProcessError("Source is not set")
        if isinstance(self._source, Jpeg):
            return self._source
        from .jpeg import decode_source  # the native decoder; raises ArtefactProcessError on failure
        return decode_source(self._source)

    def process(self) -> np.ndarray:
        """lib.rs:105-211 on the GPU; uint8 (real_px_h, real_px_w, 3)."""
        jpeg = self._decoded()
        return process_coefficients(jpeg, self._weight.to_slice(), self._pweight.to_slice(),
                                    self._iterations.to_slice(), self._separate_components, self._benchmark)


def process_coefficients(jpeg: Jpeg, weight: Sequence[float], pweight: Sequence[float],
                         iterations: Sequence[int], separate_components: bool = False,
                         benchmark: bool = False) -> np.ndarray:
    """``artefact_hip_process`` on a decoded image (three-element parameter lists)."""
    lib = _lib.load()
    if not 1 <= jpeg.nchannel <= 3 or len(jpeg.coefs) != jpeg.nchannel:
        raise _lib.ArtefactError(-1, "artefact_hip_process")
    p = _lib.artefact_process_params()
    p.nchannel = jpeg.nchannel
    p.real_w = jpeg.real_px_w
    p.real_h = jpeg.real_px_h
    for i in range(3):
        p.weight[i] = float(weight[i])
        p.pweight[i] = float(pweight[i])
        p.iterations[i] = int(iterations[i])
    p.separate_components = 1 if separate_components else 0
    p.benchmark = 1 if benchmark else 0
    chans, keep = _channels(jpeg.coefs)
    out = np.empty((jpeg.real_px_h, jpeg.real_px_w, 3), dtype=np.uint8)
    rc = lib.artefact_hip_process(ctypes.byref(p), chans, out.ctypes.data_as(_lib.U8_P))
    del keep
    if rc == _lib.ARTEFACT_BENCHMARK:
        raise ArtefactProcessError("BENCHMARK")
    _lib.check(rc, "artefact_hip_process")
    return out
