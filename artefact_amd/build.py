"""Builds libartefact_hip.so in-tree with hipcc for gfx950.

Sources: the reconstruction kernels and their C ABI (csrc/artefact_hip.hip), the warm-start
CNN kernels (csrc/cnn.hip), the host-side JPEG decoder (csrc/jpeg_decode.cpp) and image file
encoder (csrc/image_encode.cpp, plain C++ over zlib).

-ffp-contract=off keeps every f32 multiply and add separately rounded, as the
reference's default build does (backend/artefact-lib/utils/macros.rs:8-13);
packed-fp32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "artefact_hip.hip"), os.path.join(HERE, "csrc", "cnn.hip"),
           os.path.join(HERE, "csrc", "jpeg_decode.cpp"), os.path.join(HERE, "csrc", "image_encode.cpp")]
OUT = os.path.join(HERE, "libartefact_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if cand == "hipcc" or os.path.exists(cand):
            return cand
    return "hipcc"


def build_library(verbose: bool = False) -> str:
    """Compiles every HIP source into artefact_amd/libartefact_hip.so; returns its path."""
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *SOURCES, "-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
