"""Python bindings of the CPU checkers.  TEST INFRASTRUCTURE ONLY.

* ``Oracle``    -- oracle/liboracle.so, the C restatement (frehd_oracle.c);
* ``Reference`` -- oracle/_ref/libfrehd_ref.so, the reference's own step
  functions compiled from its sources with the harness ref_harness.c.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker; the product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from frehdc_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libfrehd_ref.so")

PHASE_SOURCE, PHASE_SURFACE, PHASE_UPDATE, PHASE_SCALAR, PHASE_CFL = range(5)


def build(ref: bool | None = None, quiet: bool = True):
    """Build liboracle.so, and the reference harness when the reference
    checkout is present (ref=None) or when asked (ref=True)."""
    out = None if not quiet else subprocess.DEVNULL
    subprocess.run(["make", "-C", HERE, "-s", "liboracle.so"], check=True, stdout=out)
    if ref is None:
        ref = os.path.isdir("/root/reference/FrehdC_v42")
    if ref:
        subprocess.run(["make", "-C", HERE, "-s", "ref"], check=True, stdout=out)


def _bind(lib, prefix):
    V = C.c_void_p
    getattr(lib, prefix + "create").argtypes = [C.POINTER(abi.FrehdConfig), C.POINTER(abi.FrehdForcing),
                                       C.POINTER(abi.FrehdIC)]
    getattr(lib, prefix + "create").restype = V
    getattr(lib, prefix + "destroy").argtypes = [V]
    getattr(lib, prefix + "destroy").restype = None
    getattr(lib, prefix + "phase").argtypes = [V, C.c_int]
    getattr(lib, prefix + "phase").restype = C.c_int
    getattr(lib, prefix + "step").argtypes = [V]
    getattr(lib, prefix + "step").restype = C.c_int
    getattr(lib, prefix + "field").argtypes = [V, C.c_char_p, C.POINTER(C.c_int64)]
    getattr(lib, prefix + "field").restype = C.POINTER(C.c_double)
    getattr(lib, prefix + "tt").argtypes = [V]
    getattr(lib, prefix + "tt").restype = C.c_int
    getattr(lib, prefix + "pcg_iters").argtypes = [V]
    getattr(lib, prefix + "pcg_iters").restype = C.c_int
    getattr(lib, prefix + "vloss").argtypes = [V, C.POINTER(C.c_double)]
    getattr(lib, prefix + "vloss").restype = None
    getattr(lib, prefix + "ind").argtypes = [V]
    getattr(lib, prefix + "ind").restype = C.POINTER(C.c_int32)


_libs = {}


def _lib(path, prefix):
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        _bind(lib, prefix)
        _libs[path] = lib
    return _libs[path]


class _CpuModel:
    PREFIX = ""
    PATH = ""

    def __init__(self, scen, **override):
        self.lib = _lib(self.PATH, self.PREFIX)
        self.scen = scen
        self._cfg = scen.c_config(**override)
        self._forc = scen.c_forcing()
        self._ic = scen.c_ic()
        self.h = getattr(self.lib, self.PREFIX + "create")(C.byref(self._cfg), C.byref(self._forc),
                                                  C.byref(self._ic))
        if not self.h:
            raise ValueError(f"{type(self).__name__}: configuration not supported")
        self.nScalar = max(1, self._cfg.nScalar)

    def close(self):
        if self.h:
            getattr(self.lib, self.PREFIX + "destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def phase(self, p):
        return getattr(self.lib, self.PREFIX + "phase")(self.h, p)

    def step(self, n=1):
        for _ in range(n):
            rc = getattr(self.lib, self.PREFIX + "step")(self.h)
            if rc:
                return rc
        return 0

    def raw(self, name):
        ln = C.c_int64()
        p = getattr(self.lib, self.PREFIX + "field")(self.h, name.encode(), C.byref(ln))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(ln.value,)).copy()

    def poke(self, name, values):
        """Overwrite the model's own array `name` (its first len(values)
        entries) with `values`: the models keep their state in the arrays
        the field accessor points to."""
        ln = C.c_int64()
        p = getattr(self.lib, self.PREFIX + "field")(self.h, name.encode(), C.byref(ln))
        if not p:
            raise KeyError(name)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if v.size > ln.value:
            raise ValueError(f"{name}: {v.size} values for {ln.value} slots")
        np.ctypeslib.as_array(p, shape=(ln.value,))[: v.size] = v

    @property
    def tt(self):
        return getattr(self.lib, self.PREFIX + "tt")(self.h)

    @property
    def pcg_iters(self):
        return getattr(self.lib, self.PREFIX + "pcg_iters")(self.h)

    def ind(self):
        """Level search index per cell (Ng), subgrid only."""
        p = getattr(self.lib, self.PREFIX + "ind")(self.h)
        if not p:
            raise KeyError("ind")
        return np.ctypeslib.as_array(p, shape=(self.scen.Ng,)).copy()

    def vloss(self):
        out = (C.c_double * 4)()
        getattr(self.lib, self.PREFIX + "vloss")(self.h, out)
        return np.array(out[:])


class Oracle(_CpuModel):
    PREFIX = "fo_"
    PATH = ORACLE_LIB

    def field(self, name):
        """Array in the device library's download shape: Ng for ghosted
        arrays, N for interior ones; S/Sm as (nScalar, Ng|N)."""
        s = self.scen
        a = self.raw(name)
        if name in ("S", "Sm"):
            a = a.reshape(self.nScalar, s.Ng)
            return a[:, : s.N].copy() if name == "Sm" else a
        if name in abi.INTERIOR_FIELDS:
            return a[: s.N]
        return a

    def z(self):
        return self.raw("z")[: self.scen.N]


class Reference(_CpuModel):
    PREFIX = "ref_"
    PATH = REF_LIB

    def field(self, name):
        s = self.scen
        a = self.raw(name)
        if name in ("S", "Sm"):
            return a.reshape(1, -1)
        return a

    def seam_steps(self, step_fn, nsteps):
        """SolveAll's loop body around `step_fn` (a C function with the
        oneCompleteStep signature) and the reference's monitorCFL."""
        f = self.lib.ref_seam_steps
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        f.restype = C.c_int
        return f(self.h, C.cast(step_fn, C.c_void_p), int(nsteps))

    def seam_solve_all(self, all_fn):
        """`all_fn` (a C function with SolveAll's signature) called as
        SolveAll is, on this model's state."""
        f = self.lib.ref_seam_solve_all
        f.argtypes = [C.c_void_p, C.c_void_p]
        f.restype = C.c_int
        return f(self.h, C.cast(all_fn, C.c_void_p))

    def z(self):
        """RHS in cell order (the reference stores it in the solver's
        transposed order, z[k] for cell (k % ny)*nx + k/ny)."""
        s = self.scen
        zt = self.raw("z_transposed")
        k = np.arange(s.N)
        cell = (k % s.ny) * s.nx + k // s.ny
        out = np.empty(s.N)
        out[cell] = zt
        return out


def solve(GnCt, GnXP, GnXM, GnYP, GnYM, z, eps=1e-10, max_iter=10000, precond=0):
    """The oracle's CG on the 5-point system given by row coefficients in cell
    order (nx = shape[1], ny = shape[0]); returns (x, iterations, relres)."""
    lib = _lib(ORACLE_LIB, "fo_")
    ny, nx = GnCt.shape
    P = C.POINTER(C.c_double)
    lib.fo_solve.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, P, C.c_double, C.c_int, C.c_int,
                             P]
    lib.fo_solve.restype = C.c_int
    arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
            for a in (GnCt, GnXP, GnXM, GnYP, GnYM, z)]
    x = np.zeros(nx * ny)
    rel = C.c_double()
    it = lib.fo_solve(nx, ny, *[a.ctypes.data_as(P) for a in arrs], x.ctypes.data_as(P), eps,
                      max_iter, precond, C.byref(rel))
    return x.reshape(ny, nx), it, rel.value


def have_reference():
    return os.path.exists(REF_LIB)
