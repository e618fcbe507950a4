"""ctypes mirror of include/frehd_gpu.h and the loader of the device library.

The structs below must match the C declarations field for field; the test
``tests/test_abi.py`` checks their sizes against the compiled library.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfrehd_gpu.so")

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)


class FrehdConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("Nt", C.c_int32), ("maxIter", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dt", C.c_double), ("g", C.c_double),
        ("NUx", C.c_double), ("NUy", C.c_double), ("CDx", C.c_double), ("CDy", C.c_double),
        ("manningN", C.c_double),
        ("CDnotN", C.c_int32), ("useThinLayer", C.c_int32),
        ("hD", C.c_double), ("CDmax", C.c_double), ("minDepth", C.c_double),
        ("wtfh", C.c_double), ("CFLl", C.c_double), ("CFLh", C.c_double), ("eps", C.c_double),
        ("precond", C.c_int32), ("precision", C.c_int32), ("bcType", C.c_int32),
        ("tideLocLengthP", C.c_int32), ("tideLocLengthM", C.c_int32),
        ("inflowLocLength", C.c_int32), ("useScalar", C.c_int32), ("scalarAdv", C.c_int32),
        ("useConstTidePS", C.c_int32), ("useConstTideMS", C.c_int32), ("nScalar", C.c_int32),
        ("tidePS", C.c_double), ("tideMS", C.c_double), ("inflowS", C.c_double),
        ("Kx", C.c_double), ("Ky", C.c_double), ("offset", C.c_double),
        ("epsInner", C.c_double), ("maxRefine", C.c_int32),
        ("useSubgrid", C.c_int32), ("Ne", C.c_int32), ("warmStart", C.c_int32),
        ("surfmin", C.c_double), ("dsurf", C.c_double), ("subR", C.c_double),
        ("subgridOnTheFly", C.c_int32), ("useWind", C.c_int32),
        ("northAngle", C.c_double), ("rhoa", C.c_double), ("Cw", C.c_double), ("CwT", C.c_double),
        ("useSubDrag", C.c_int32), ("useminA", C.c_int32), ("useminV", C.c_int32), ("pad3", C.c_int32),
        ("beta", C.c_double),
    ]


class FrehdForcing(C.Structure):
    _fields_ = [
        ("tideLocP", c_int32_p), ("tideLocM", c_int32_p), ("inflowLoc", c_int32_p),
        ("tideP", c_double_p), ("tideM", c_double_p), ("inflow", c_double_p),
        ("tidalPS", c_double_p), ("tidalMS", c_double_p), ("manningCell", c_double_p),
        ("windspd", c_double_p), ("winddir", c_double_p),
    ]


SUB_TABLES = ["V", "Z", "Vxp", "Vxm", "Vyp", "Vym", "Np", "Nm", "Op", "Om",
              "minVxp", "minVxm", "minVyp", "minVym"]
DRAG_TABLES = ["YhX", "YhY", "Cn"]
MINA_TABLES = ["Nmin", "Nmax", "Omin", "Omax"]


class FrehdSubgrid(C.Structure):
    _fields_ = ([("edgeX", c_double_p), ("edgeY", c_double_p)] + [(n, c_double_p) for n in SUB_TABLES]
                + [("fine", c_double_p), ("fineR", C.c_int32), ("pad", C.c_int32), ("dxf", C.c_double)]
                + [(n, c_double_p) for n in DRAG_TABLES] + [(n, c_double_p) for n in MINA_TABLES])


class FrehdIC(C.Structure):
    _fields_ = [
        ("bottomZ", c_double_p),
        ("surf0", C.c_double), ("U0", C.c_double), ("V0", C.c_double), ("S0", C.c_double),
        ("surf0cell", c_double_p), ("U0cell", c_double_p), ("V0cell", c_double_p),
        ("S0cell", c_double_p), ("sub", C.POINTER(FrehdSubgrid)),
    ]


FIELD_NAMES = [
    "surf", "surfOld", "uuXP", "vvYP", "uuYP", "vvXP", "depth", "depthXP", "depthYP",
    "EnXP", "EnYP", "DragXP", "DragYP", "CDXP", "CDYP", "Fuu", "Fvv", "cellV", "S", "Sm",
    "CFLx", "CFLy", "wtfXP", "wtfYP", "bottomZ", "bottomXP", "bottomYP",
]
# arrays that hold only the N interior values (the rest hold Ng)
INTERIOR_FIELDS = {"uuYP", "vvXP", "depth", "cellV", "Sm"}


SUB_FIELD_NAMES = ["V", "Z", "Vx", "Vy", "Zx", "Zy", "Nx", "Oy"]


class FrehdFields(C.Structure):
    _fields_ = [(n, c_double_p) for n in FIELD_NAMES + SUB_FIELD_NAMES] + [("ind", c_int32_p)]


class FrehdStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("pcgIterTotal", C.c_int64), ("pcgIterLast", C.c_int32),
        ("refineLast", C.c_int32), ("relResLast", C.c_double), ("maxCFLx", C.c_double),
        ("maxCFLy", C.c_double), ("Vloss", C.c_double * 4), ("halted", C.c_int32),
        ("haltStep", C.c_int32),
    ]


MAX_SCALARS = 16


class FrehdBudget(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("nScalar", C.c_int32), ("on", C.c_int32), ("volume", C.c_double),
        ("volumeResidual", C.c_double), ("volumeResidualMax", C.c_double), ("volumeBoundary", C.c_double),
        ("volumeInflow", C.c_double), ("volumeClip", C.c_double), ("volumeMoved", C.c_double),
        ("mass", C.c_double * MAX_SCALARS), ("massResidual", C.c_double * MAX_SCALARS),
        ("massResidualMax", C.c_double * MAX_SCALARS), ("massBoundary", C.c_double * MAX_SCALARS),
        ("massInflow", C.c_double * MAX_SCALARS), ("massMoved", C.c_double * MAX_SCALARS),
    ]


FREHD_OK = 0
FREHD_EINVAL = -1
FREHD_EHIP = -2
FREHD_ENOMEM = -3
FREHD_ECFL = -4
FREHD_EBUDGET = -5
FREHD_ESTATE = -6

PRECOND_JACOBI = 0
PRECOND_SSOR = 1
PREC_FP64 = 0
PREC_MIXED = 1


def dptr(a):
    """double* to a contiguous float64 numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_double_p)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_int32_p)


_lib = None


class FrehdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"frehd error {code}: {msg}")
        self.code = code


class FrehdSubgridSpec(C.Structure):
    """include/frehd_subgrid.h frehd_subgrid_spec"""
    _fields_ = [("nxFine", C.c_int32), ("nyFine", C.c_int32), ("r", C.c_int32), ("Ne", C.c_int32),
                ("dxf", C.c_double), ("surfmin", C.c_double), ("dsurf", C.c_double),
                ("checkBlock", C.c_int32), ("useminVolume", C.c_int32), ("useYh", C.c_int32),
                ("useCn", C.c_int32), ("n0", C.c_double)]


SUBGRID_TABLES = ("V", "Z", "Vxp", "Vxm", "Vyp", "Vym", "Np", "Nm", "Op", "Om", "Nmin", "Nmax", "Omin", "Omax",
                  "minVxp", "minVxm", "minVyp", "minVym", "YhX", "YhY", "Cn")
SUBGRID_GEOMETRY = ("bottom", "bottomXP", "bottomYP")


class FrehdSubgridOut(C.Structure):
    """include/frehd_subgrid.h frehd_subgrid_out"""
    _fields_ = [(n, c_double_p) for n in SUBGRID_TABLES + SUBGRID_GEOMETRY]


def load_library(path: str | None = None):
    """Load libfrehd_gpu.so (built in-tree).  Raises if it is missing: the
    device path has no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(p)
    V = C.c_void_p
    sig = {
        "frehd_last_error": ([], C.c_char_p),
        "frehd_version": ([], C.c_char_p),
        "frehd_device_count": ([c_int32_p], C.c_int),
        "frehd_dev_create": ([C.POINTER(FrehdConfig), C.POINTER(FrehdForcing),
                              C.POINTER(FrehdIC), C.POINTER(V)], C.c_int),
        "frehd_dev_create_from_state": ([C.POINTER(FrehdConfig), C.POINTER(FrehdForcing),
                                         C.POINTER(FrehdSubgrid), C.POINTER(FrehdFields),
                                         C.POINTER(V)], C.c_int),
        "frehd_dev_step": ([V, C.c_int32], C.c_int),
        "frehd_dev_step_noalloc": ([V, C.c_int32], C.c_int),
        "frehd_dev_download": ([V, C.POINTER(FrehdFields)], C.c_int),
        "frehd_dev_stats": ([V, C.POINTER(FrehdStats)], C.c_int),
        "frehd_dev_set_budget": ([V, C.c_int32], C.c_int),
        "frehd_dev_budget": ([V, C.POINTER(FrehdBudget)], C.c_int),
        "frehd_dev_step_index": ([V, c_int32_p], C.c_int),
        "frehd_dev_set_step_index": ([V, C.c_int32], C.c_int),
        "frehd_dev_phase": ([V, C.c_int32], C.c_int),
        "frehd_dev_destroy": ([V], C.c_int),
        "frehd_dev_footprint": ([C.POINTER(FrehdConfig), C.POINTER(C.c_int64)], C.c_int),
        "frehd_dev_autotune": ([V, C.c_int32, c_int32_p, c_int32_p, c_int32_p, C.POINTER(C.c_double)], C.c_int),
        "frehd_batch_autotune": ([V, C.c_int32, c_int32_p, c_int32_p, c_int32_p, C.POINTER(C.c_double)], C.c_int),
        "frehd_dev_enable_timing": ([V, C.c_int32], C.c_int),
        "frehd_dev_kernel_times": ([V, C.c_int32, C.POINTER(C.c_char_p), c_double_p,
                                    C.POINTER(C.c_int64), c_int32_p], C.c_int),
        "frehd_abi_sizes": ([C.POINTER(C.c_int64)], C.c_int),
        "frehd_dev_set_pcg_tiling": ([V, C.c_int32, C.c_int32, C.c_int32], C.c_int),
        "frehd_dev_set_fused_solve": ([V, C.c_int32], C.c_int),
        "frehd_dev_autotune_blocks": ([V, C.c_int32, c_int32_p, C.POINTER(C.c_double)], C.c_int),
        "frehd_batch_autotune_blocks": ([V, C.c_int32, c_int32_p, C.POINTER(C.c_double)], C.c_int),
        "frehd_batch_set_fused_solve": ([V, C.c_int32], C.c_int),
        "frehd_fused_solve_capacity": ([C.c_int32, c_int32_p], C.c_int),
        "frehd_dev_upload": ([V, C.POINTER(FrehdFields)], C.c_int),
        "frehd_dev_download_tables": ([V, C.POINTER(c_double_p)], C.c_int),
        "frehd_batch_create": ([C.c_int32, C.POINTER(V)], C.c_int),
        "frehd_batch_add": ([V, V], C.c_int),
        "frehd_batch_remove": ([V, V], C.c_int),
        "frehd_batch_size": ([V, c_int32_p], C.c_int),
        "frehd_batch_step": ([V, C.c_int32], C.c_int),
        "frehd_batch_enable_timing": ([V, C.c_int32], C.c_int),
        "frehd_batch_kernel_times": ([V, C.c_int32, C.POINTER(C.c_char_p), c_double_p,
                                      C.POINTER(C.c_int64), c_int32_p], C.c_int),
        "frehd_batch_destroy": ([V], C.c_int),
        "frehd_subgrid_generate": ([C.POINTER(C.c_double), C.POINTER(FrehdSubgridSpec),
                                    C.POINTER(FrehdSubgridOut), C.c_int32], C.c_int),
        "frehd_subgrid_write_text": ([C.c_char_p, C.POINTER(FrehdSubgridSpec), C.POINTER(FrehdSubgridOut)],
                                     C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc: int):
    if rc != 0:
        msg = _lib.frehd_last_error().decode() if _lib is not None else ""
        raise FrehdError(rc, msg)
    return rc
