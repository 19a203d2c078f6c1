"""MI355X-native reconstruction loop of artefact-lib (HIP kernels behind a C ABI).

``compute`` mirrors backend/artefact-lib/pipeline_scalar/mod.rs:17-26; ``Plan`` is
its resident form.  ``Coefficient`` mirrors backend/artefact-lib/jpeg/mod.rs:8-27.
``Artefact`` mirrors the public builder of backend/artefact-lib/lib.rs:49-212.
"""
from ._lib import ArtefactError, LIB_PATH
from .batch import Batch, BatchImage
from .compute import Plan, compute
from .process import Artefact, ArtefactProcessError, ForAll, ForEach, Jpeg, JpegSource, process_coefficients
from .jpeg import JpegDecodeError, decode as decode_jpeg
from .image import ImageFormatError, encode as encode_image, save as save_image
from .types import Coefficient, max_grid

__all__ = ["Batch", "BatchImage", "Artefact", "ArtefactError", "ArtefactProcessError", "Coefficient", "ForAll", "ForEach", "Jpeg",
           "JpegSource", "JpegDecodeError", "decode_jpeg", "ImageFormatError", "encode_image", "save_image", "LIB_PATH", "Plan", "compute", "max_grid", "process_coefficients"]
