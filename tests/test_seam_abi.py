"""The reference-signature entry points (frehdc_amd/seam/frehd_seam.c,
include/frehd_seam.h) are compiled against FrehdC's own headers, so the
compiler holds oneCompleteStep to nssolve.h:10-11; the library exports it and
every function include/frehd_seam.h declares.  No device call is made here
(the entry points exit the process on a failure, as the reference does)."""
import os
import re
import subprocess

import pytest

from frehdc_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "frehd_seam.h")


def declared():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\bvoid\s+(\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def seam_lib():
    if os.path.isfile(os.path.join(build.REF_HEADERS, "nssolve.h")):
        if not os.path.exists(build.OUT):
            build.build()
        else:
            build.build_seam()
    if not os.path.exists(build.SEAM_LIB):
        pytest.skip("no FrehdC headers here and no prebuilt libfrehd_seam.so")
    return build.SEAM_LIB


def test_seam_exports(seam_lib):
    names = declared() + ["oneCompleteStep"]
    assert {"SolveAllOnDevice", "frehd_seam_release"} <= set(names)
    out = subprocess.run(["nm", "-D", "--defined-only", seam_lib], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(\w+)", out))
    assert not [n for n in names if n not in exported]


def test_seam_links_the_device_library(seam_lib):
    out = subprocess.run(["readelf", "-d", seam_lib], capture_output=True, text=True, check=True).stdout
    assert "libfrehd_gpu.so" in out and "$ORIGIN/.." in out
    # the entry points need nothing of the reference but its output routines
    und = subprocess.run(["nm", "-D", "--undefined-only", seam_lib], capture_output=True, text=True,
                         check=True).stdout
    ref_syms = {s for s in re.findall(r"\sU\s(\w+)", und) if not s.startswith(("frehd_", "__"))}
    libc = {"printf", "exit", "calloc", "free", "memset", "memcpy", "puts", "putchar", "getenv", "strcmp"}
    assert ref_syms - libc == {"DataOutput", "SubOutput"}, ref_syms
