"""ctypes binding of the C ABI declared in include/artefact_hip.h.

The shared library is built in-tree (``artefact_amd/libartefact_hip.so``).  There
is no CPU fallback: when the library is missing or cannot be loaded, every entry
point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libartefact_hip.so")

ARTEFACT_MODE_TGV = 0
ARTEFACT_MODE_TV = 1
ARTEFACT_PRECISION_F32 = 0
ARTEFACT_PRECISION_BF16X3 = 1
ARTEFACT_PRECISION_F32_FAST = 2

ERRORS = {
    -1: "invalid argument",
    -2: "invalid geometry",
    -3: "zero entry in a quantisation table",
    -4: "HIP runtime failure",
    -5: "plan used out of order",
    -6: "malformed or unsupported JPEG data",
    -7: "grid too large for the no-allocation entry point's static scratch",
    1: "benchmark run: the loop ran and no image was written",
}
ARTEFACT_BENCHMARK = 1
ARTEFACT_E_SCRATCH = -7
ARTEFACT_IMAGE_PNG = 0
ARTEFACT_IMAGE_BMP = 1
ARTEFACT_IMAGE_TIFF = 2
ARTEFACT_IMAGE_WEBP = 3
ARTEFACT_SCRATCH_BYTES = 64 << 20


class ArtefactError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"{where}: {ERRORS.get(code, 'unknown error')} (code {code})")
        self.code = code


class artefact_channel(ctypes.Structure):
    _fields_ = [
        ("rounded_w", ctypes.c_uint32),
        ("rounded_h", ctypes.c_uint32),
        ("block_w", ctypes.c_uint32),
        ("block_h", ctypes.c_uint32),
        ("h_samp", ctypes.c_uint32),
        ("v_samp", ctypes.c_uint32),
        ("coefs", ctypes.POINTER(ctypes.c_int16)),
        ("quant", ctypes.POINTER(ctypes.c_uint16)),
    ]


class artefact_params(ctypes.Structure):
    _fields_ = [
        ("nchannel", ctypes.c_uint32),
        ("max_w", ctypes.c_uint32),
        ("max_h", ctypes.c_uint32),
        ("weight", ctypes.c_float),
        ("pweight", ctypes.c_float * 3),
        ("iterations", ctypes.c_uint32),
        ("mode", ctypes.c_uint32),
        ("init_override", ctypes.POINTER(ctypes.POINTER(ctypes.c_float))),
    ]


class artefact_process_params(ctypes.Structure):
    _fields_ = [
        ("nchannel", ctypes.c_uint32),
        ("real_w", ctypes.c_uint32),
        ("real_h", ctypes.c_uint32),
        ("weight", ctypes.c_float * 3),
        ("pweight", ctypes.c_float * 3),
        ("iterations", ctypes.c_uint32 * 3),
        ("separate_components", ctypes.c_uint32),
        ("benchmark", ctypes.c_uint32),
    ]


class artefact_tuning(ctypes.Structure):
    _fields_ = [
        ("kg_rows", ctypes.c_uint32),
        ("project_kernel", ctypes.c_uint32),
        ("ms_per_iteration", ctypes.c_float),
        ("project_workgroups", ctypes.c_uint32),
        ("xcd_bands", ctypes.c_uint32),
    ]


class artefact_batch_image(ctypes.Structure):
    _fields_ = [
        ("params", ctypes.POINTER(artefact_params)),
        ("channels", ctypes.POINTER(artefact_channel)),
    ]


FLOAT_PP = ctypes.POINTER(ctypes.POINTER(ctypes.c_float))
U8_P = ctypes.POINTER(ctypes.c_uint8)

# (name, restype, argtypes) of every exported symbol of include/artefact_hip.h.
SIGNATURES = [
    ("artefact_hip_compute", ctypes.c_int, [ctypes.POINTER(artefact_params), ctypes.POINTER(artefact_channel), FLOAT_PP]),
    ("artefact_image_encode_bound", ctypes.c_size_t, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    ("artefact_image_encode", ctypes.c_int,
     [U8_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, U8_P, ctypes.c_size_t,
      ctypes.POINTER(ctypes.c_size_t)]),
    ("artefact_hip_compute_noalloc", ctypes.c_int,
     [ctypes.POINTER(artefact_params), ctypes.POINTER(artefact_channel), FLOAT_PP, ctypes.c_void_p]),
    ("artefact_hip_plan_create", ctypes.c_int,
     [ctypes.POINTER(artefact_params), ctypes.POINTER(artefact_channel), ctypes.POINTER(ctypes.c_void_p)]),
    ("artefact_hip_plan_init", ctypes.c_int, [ctypes.c_void_p]),
    ("artefact_hip_plan_iterate", ctypes.c_int, [ctypes.c_void_p]),
    ("artefact_hip_plan_download", ctypes.c_int, [ctypes.c_void_p, FLOAT_PP]),
    ("artefact_hip_plan_download_cos", ctypes.c_int, [ctypes.c_void_p, FLOAT_PP]),
    ("artefact_hip_plan_last_norms", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("artefact_hip_plan_first_gradient", ctypes.c_int, [ctypes.c_void_p, FLOAT_PP]),
    ("artefact_hip_plan_profile", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("artefact_hip_plan_set_timing", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("artefact_hip_plan_timing", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint32)]),
    ("artefact_hip_plan_objective", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    ("artefact_hip_plan_stream", ctypes.c_void_p, [ctypes.c_void_p]),
    ("artefact_hip_plan_set_graph", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("artefact_hip_plan_destroy", None, [ctypes.c_void_p]),
    ("artefact_hip_plan_download_rgb", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, U8_P]),
    ("artefact_hip_process", ctypes.c_int,
     [ctypes.POINTER(artefact_process_params), ctypes.POINTER(artefact_channel), U8_P]),
    ("artefact_hip_plan_tune", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(artefact_tuning)]),
    ("artefact_hip_plan_set_precision", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("artefact_hip_batch_create", ctypes.c_int,
     [ctypes.POINTER(artefact_batch_image), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
      ctypes.POINTER(ctypes.c_void_p)]),
    ("artefact_hip_batch_run", ctypes.c_int, [ctypes.c_void_p]),
    ("artefact_hip_batch_streams", ctypes.c_uint32, [ctypes.c_void_p]),
    ("artefact_hip_batch_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("artefact_hip_batch_download", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, FLOAT_PP]),
    ("artefact_hip_batch_plan", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]),
    ("artefact_hip_batch_destroy", None, [ctypes.c_void_p]),
    ("artefact_cnn_param_count", ctypes.c_uint32, []),
    ("artefact_cnn_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_void_p)]),
    ("artefact_cnn_get_params", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("artefact_cnn_forward", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_uint32,
      ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)]),
    ("artefact_cnn_gradients", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double),
      ctypes.POINTER(ctypes.c_float)]),
    ("artefact_cnn_set_dataset", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
      ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    ("artefact_cnn_train_step", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.c_float,
      ctypes.POINTER(ctypes.c_double)]),
    ("artefact_cnn_destroy", None, [ctypes.c_void_p]),
    ("artefact_hip_plan_warm_init", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("artefact_hip_plan_set_step_iterations", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("artefact_hip_plan_set_momentum_offset", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("artefact_hip_plan_set_step_schedule", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    ("artefact_jpeg_decode", ctypes.c_int,
     [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p), ctypes.c_char_p, ctypes.c_size_t]),
    ("artefact_jpeg_nchannel", ctypes.c_uint32, [ctypes.c_void_p]),
    ("artefact_jpeg_size", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    ("artefact_jpeg_channels", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(artefact_channel)]),
    ("artefact_jpeg_free", None, [ctypes.c_void_p]),
    ("artefact_hip_strerror", ctypes.c_char_p, [ctypes.c_int]),
]

_lib = None


def load() -> ctypes.CDLL:
    """Loads the HIP library once; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, restype, argtypes in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def check(code: int, where: str) -> None:
    if code != 0:
        raise ArtefactError(code, where)
