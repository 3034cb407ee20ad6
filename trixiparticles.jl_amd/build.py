"""Build step of the package: compiles the HIP library ``libtpsph.so`` for gfx950 in-tree.

Run as ``python trixiparticles.jl_amd/build.py`` or call :func:`build_library`."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libtpsph.so")
SOURCES = [os.path.join(PKG_DIR, "csrc", f) for f in ("tpsph.hip", "tpsph_mixed.hip", "tpsph_batch.hip", "tpsph_adjoint.hip")]
HEADERS = [os.path.join(PKG_DIR, "csrc", "tpsph_internal.h")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # no packed FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc_path() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build_library(verbose: bool = False) -> str:
    """Compile ``libtpsph.so`` (always recompiles) and return its path."""
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc_path(), *HIPCC_FLAGS, "-I", os.path.join(REPO_DIR, "include"),
           "-o", tmp, *SOURCES]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose:
        noise = "packed-fp32-ops"
        msg = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if noise not in l)
        if msg.strip():
            print(msg, file=sys.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
