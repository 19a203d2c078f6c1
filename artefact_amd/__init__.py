"""MI355X-native reconstruction loop of artefact-lib (HIP kernels behind a C ABI).

``compute`` mirrors backend/artefact-lib/pipeline_scalar/mod.rs:17-26; ``Plan`` is
its resident form.  ``Coefficient`` mirrors backend/artefact-lib/jpeg/mod.rs:8-27.
"""
from ._lib import ArtefactError, LIB_PATH
from .compute import Plan, compute
from .types import Coefficient, max_grid

__all__ = ["ArtefactError", "Coefficient", "LIB_PATH", "Plan", "compute", "max_grid"]
