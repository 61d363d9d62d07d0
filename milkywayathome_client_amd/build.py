"""Build the nbhip HIP library in-tree for gfx950.

    python -m milkywayathome_client_amd.build

produces ``milkywayathome_client_amd/libnbhip.so``.  The build always
recompiles (it is the "does it build" check); loaders never build.
"""
from __future__ import annotations

import glob
import hashlib
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libnbhip.so")
MODELS_LIB_PATH = os.path.join(PKG_DIR, "libnbmodels.so")
# Digest of the sources and flags the two libraries were built from, written
# beside them by build() and checked by the loader (a library built from
# other sources is refused, not used).
STAMP_PATH = os.path.join(PKG_DIR, "libnbhip.srchash")

SOURCES = ["api.hip", "tree.hip", "walk.hip", "walk_group.hip", "adjoint.hip", "histogram.hip",
           "exact.hip", "run.hip", "checkpoint.hip", "dwarf.hip"]

# -ffp-contract=off: the plain FP64 form keeps the reference's two-rounding
# a*b+c so that it reproduces the CPU path bit for bit.
# -packed-fp32-ops: no v_pk_* FP32 instructions are generated.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def source_hash() -> str:
    """SHA-256 over every source file, header and compiler flag of the build."""
    h = hashlib.sha256()
    files = sorted(glob.glob(os.path.join(CSRC, "*")) +
                   glob.glob(os.path.join(PKG_DIR, "..", "include", "*.h")))
    for f in files:
        if os.path.isfile(f):
            h.update(os.path.basename(f).encode())
            with open(f, "rb") as fh:
                h.update(fh.read())
    h.update(" ".join(HIPCC_FLAGS + SOURCES).encode())
    return h.hexdigest()


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the nbhip library needs ROCm's hipcc to build")


def build_models(verbose: bool = False) -> str:
    """Host-only initial-condition library (plain C, gcc)."""
    cc = os.environ.get("CC") or shutil.which("gcc") or "cc"
    tmp = MODELS_LIB_PATH + ".tmp"
    cmd = [cc, "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
           os.path.join(CSRC, "models.c"), "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("gcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, MODELS_LIB_PATH)
    return MODELS_LIB_PATH


def build(verbose: bool = False) -> str:
    build_models(verbose)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    cmd = [hipcc_path(), *HIPCC_FLAGS, "-I", os.path.join(PKG_DIR, "..", "include"),
           *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB_PATH)
    with open(STAMP_PATH, "w") as f:
        f.write(source_hash() + "\n")
    return LIB_PATH


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
