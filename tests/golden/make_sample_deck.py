"""Copy the reference's shipped sample input deck (FrehdC_v42/data.dat and
its plain-grid input files) into tests/golden/sample_deck/ and record what
the reference's own readers make of it (oracle/_ref: readOneBC for the
boundary series, ReadBathymetry + InitBathymetry for the bed and offset)
in tests/golden/sample_deck/reference_readers.npz.

    TZ=CST6 python tests/golden/make_sample_deck.py

TZ=CST6 makes the deck's start date (2013-06-10, local midnight) the first
time stamp of its tide files, as the shipped deck assumes.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = "/root/reference/FrehdC_v42"
DST = os.path.join(HERE, "sample_deck")
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import pyoracle  # noqa: E402


def main():
    os.environ.setdefault("TZ", "CST6")
    time.tzset()
    os.makedirs(os.path.join(DST, "input"), exist_ok=True)
    shutil.copyfile(os.path.join(SRC, "data.dat"), os.path.join(DST, "data.dat"))
    for f in ("bath.dat", "tideP.dat", "tideM.dat", "inflow.dat"):
        shutil.copyfile(os.path.join(SRC, "input", f), os.path.join(DST, "input", f))
    lib = C.CDLL(pyoracle.REF_LIB)
    P = C.POINTER(C.c_double)
    tN = 1370844000.0          # 2013-06-10 00:00 CST6
    dt, Nt = 2.0, 43200
    out = {}
    cwd = os.getcwd()
    os.chdir(DST)              # the reference builds paths in a 50-byte buffer
    try:
        for name, N in (("tideP.dat", 721), ("inflow.dat", 487)):
            a = np.zeros(Nt)
            lib.ref_io_read_bc(b"input/", name.encode(), N, C.c_double(tN), C.c_double(dt), Nt, a.ctypes.data_as(P))
            out[name.split(".")[0]] = a
        z = np.zeros(30 * 124)
        off = C.c_double()
        lib.ref_io_bathymetry(b"input/", b"", 0, 30, 124, z.ctypes.data_as(P), None, None, C.byref(off))
        out["allZ"] = z
        out["offset"] = np.array([off.value])
    finally:
        os.chdir(cwd)
    np.savez_compressed(os.path.join(DST, "reference_readers.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
