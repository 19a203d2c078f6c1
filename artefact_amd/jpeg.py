"""JPEG input stage: ``Jpeg::from`` (backend/artefact-lib/jpeg/zune.rs:5-62).

``decode(data)`` runs the library's host-side decoder (csrc/jpeg_decode.cpp) and returns
a ``Jpeg`` whose ``Coefficient``s are laid out as the zune-jpeg fork lays them out.
``decode_source`` takes a ``JpegSource`` and raises ``ArtefactProcessError`` with the
reference's messages ("Failed to read JPEG file '<path>': ...", "Failed to decode JPEG:
...", lib.rs:99-103 with zune.rs:7-15).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .process import ArtefactProcessError, Jpeg, JpegSource
from .types import Coefficient


class JpegDecodeError(ValueError):
    pass


def decode(data: bytes) -> Jpeg:
    """Coefficients of a baseline or progressive JPEG held in memory."""
    lib = _lib.load()
    data = bytes(data)
    h = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    rc = lib.artefact_jpeg_decode(data, len(data), ctypes.byref(h), err, len(err))
    if rc == -6:
        raise JpegDecodeError(err.value.decode(errors="replace"))
    _lib.check(rc, "artefact_jpeg_decode")
    try:
        n = lib.artefact_jpeg_nchannel(h)
        w, hh = ctypes.c_uint32(), ctypes.c_uint32()
        _lib.check(lib.artefact_jpeg_size(h, ctypes.byref(w), ctypes.byref(hh)), "artefact_jpeg_size")
        chans = (_lib.artefact_channel * n)()
        _lib.check(lib.artefact_jpeg_channels(h, chans), "artefact_jpeg_channels")
        coefs = []
        for c in chans:
            nb = c.block_w * c.block_h * 64
            d = np.ctypeslib.as_array(c.coefs, shape=(nb,)).copy()
            q = np.ctypeslib.as_array(c.quant, shape=(64,)).copy()
            coefs.append(Coefficient(rounded_px_w=c.rounded_w, rounded_px_h=c.rounded_h, block_w=c.block_w,
                                     block_h=c.block_h, horizontal_samp_factor=c.h_samp,
                                     vertical_samp_factor=c.v_samp, dct_coefs=d, quant_table=q))
    finally:
        lib.artefact_jpeg_free(h)
    return Jpeg(nchannel=n, real_px_w=w.value, real_px_h=hh.value, coefs=coefs)


def decode_source(source: JpegSource) -> Jpeg:
    if source.path is not None:
        try:
            with open(source.path, "rb") as fh:
                data = fh.read()
        except OSError as e:
            raise ArtefactProcessError(f"Failed to read JPEG: Failed to read JPEG file '{source.path}': {e}") from e
    else:
        data = source.buffer or b""
    try:
        return decode(data)
    except JpegDecodeError as e:
        raise ArtefactProcessError(f"Failed to read JPEG: Failed to decode JPEG: {e}") from e
