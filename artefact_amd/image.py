"""RGB image -> file: the save step of the reference's front ends.

``Artefact().process()`` returns the RGB image; the reference's CLI then picks the output
path and format (backend/artefact-cli/main.rs:57-98) and saves with the ``image`` crate
(``img.save(output)``, main.rs:156), which chooses the encoder from the file extension.
``output_path`` restates that choice, ``save`` the save, ``encode`` the encoder call
(csrc/image_encode.cpp: PNG, BMP, TIFF and lossless WebP; GIF, which the CLI also names, is
not built and raises).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib

FORMATS = {"png": _lib.ARTEFACT_IMAGE_PNG, "bmp": _lib.ARTEFACT_IMAGE_BMP, "tiff": _lib.ARTEFACT_IMAGE_TIFF,
           "webp": _lib.ARTEFACT_IMAGE_WEBP}
# The CLI's list (main.rs:52) and the extensions the image crate maps to a format.
POSSIBLE_FORMATS = ["png", "webp", "tiff", "bmp", "gif"]
EXTENSIONS = {"png": "png", "bmp": "bmp", "tif": "tiff", "tiff": "tiff", "webp": "webp", "gif": "gif"}


class ImageFormatError(ValueError):
    """An output format the CLI rejects, or one this build does not encode."""


def encode(rgb: np.ndarray, format: str = "png") -> bytes:
    """File bytes of a uint8 (height, width, 3) RGB image in ``format`` (png, bmp, tiff or webp)."""
    fmt = FORMATS.get(format.lower())
    if fmt is None:
        raise ImageFormatError(f"cannot encode {format!r}: this build writes {', '.join(FORMATS)}")
    img = np.ascontiguousarray(rgb, dtype=np.uint8)
    if img.ndim != 3 or img.shape[2] != 3 or img.shape[0] == 0 or img.shape[1] == 0:
        raise ValueError("expected a non-empty (height, width, 3) uint8 image")
    h, w = img.shape[:2]
    lib = _lib.load()
    cap = lib.artefact_image_encode_bound(w, h, fmt)
    out = np.empty(cap, dtype=np.uint8)
    n = ctypes.c_size_t()
    _lib.check(lib.artefact_image_encode(img.ctypes.data_as(_lib.U8_P), w, h, fmt, out.ctypes.data_as(_lib.U8_P),
                                         cap, ctypes.byref(n)), "artefact_image_encode")
    return out[: n.value].tobytes()


def output_path(input_path: str, output: str | None = None, format: str = "auto") -> str:
    """The CLI's output path (main.rs:57-98): with format "auto" the output's extension, or png;
    otherwise format must be one of POSSIBLE_FORMATS and match the output's extension; an output
    without extension gets the format's, and no output means the input path with it."""
    if format == "auto":
        final = (os.path.splitext(output)[1][1:] or "png") if output is not None else "png"
    else:
        if format not in POSSIBLE_FORMATS:
            raise ImageFormatError(f"Invalid output format ({format}), possible values: {', '.join(POSSIBLE_FORMATS)}")
        final = format
    if output is not None:
        root, ext = os.path.splitext(output)
        if ext:
            if format != "auto" and ext[1:] != final:
                raise ImageFormatError("Output file extension does not match output format")
            return output
        return f"{output}.{final}"
    return f"{os.path.splitext(input_path)[0]}.{final}"


def save(rgb: np.ndarray, path: str) -> None:
    """Writes the image in the format its extension names (png, bmp, tif / tiff, webp)."""
    ext = os.path.splitext(path)[1][1:].lower()
    fmt = EXTENSIONS.get(ext)
    if fmt is None:
        raise ImageFormatError(f"The image format could not be determined from the extension {ext!r}")
    data = encode(rgb, fmt)
    with open(path, "wb") as f:
        f.write(data)
