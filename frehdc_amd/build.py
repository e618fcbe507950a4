"""Build the device library libfrehd_gpu.so in-tree with hipcc for gfx950, and
the reference-signature entry points beside it (seam/libfrehd_seam.so) where
the FrehdC headers are present."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfrehd_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["api.hip", "exec.hip", "step_kernels.hip", "pcg_kernels.hip", "subgrid_tables.hip", "noalloc.hip",
           "case_io.cpp", "subgrid_gen.cpp"]

# -ffp-contract=off: no fused multiply-add, so every FP64 expression rounds as
# the reference's (x86-64, no FMA) does.  Packed FP32 instructions
# (v_pk_*_f32) are disabled as a precaution.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False, out: str = OUT, defines=()) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-I", os.path.join(os.path.dirname(HERE), "include"),
           "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libfrehd_gpu.so")
    os.replace(tmp, out)
    if out == OUT:
        build_seam(verbose)
    return out


SEAM_DIR = os.path.join(HERE, "seam")
SEAM_LIB = os.path.join(SEAM_DIR, "libfrehd_seam.so")
REF_HEADERS = os.environ.get("FREHD_REFERENCE", "/root/reference") + "/FrehdC_v42"


def build_seam(verbose: bool = False) -> str | None:
    """The reference-signature entry points (seam/frehd_seam.c), compiled
    against the FrehdC headers where a checkout is present; elsewhere the
    library built next to one travels prebuilt."""
    if not os.path.isfile(os.path.join(REF_HEADERS, "nssolve.h")):
        return None
    cmd = ["make", "-C", SEAM_DIR, "-B", "-s", f"REFROOT={os.path.dirname(REF_HEADERS)}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building libfrehd_seam.so failed")
    return SEAM_LIB


if __name__ == "__main__":
    print(build(verbose=True))
