"""FrehdC input decks on the device library (include/frehd_io.h).

    case = Case("data.dat", base_dir)          # data.dat, bathymetry, BCs, ICs, subgrid tables
    scen = case.scenario()                     # the same inputs as a cases.Scenario
    run_case("data.dat", base_dir, steps=900)  # the reference's time loop, outputs included

`run_case` is SolveAll (FrehdC_v42/nssolve.c:48-106) on the device: output of
the initial state (step 0), then every step with the CFL stop test, and the
DataOutput / SubOutput files every OutItvl steps.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi, cases

V = C.c_void_p


class CaseInfo(C.Structure):
    _fields_ = [
        ("Nt", C.c_int32), ("OutItvl", C.c_int32), ("savesurface", C.c_int32), ("saveuu", C.c_int32),
        ("savevv", C.c_int32), ("savedepth", C.c_int32), ("savescalar", C.c_int32), ("saveCD", C.c_int32),
        ("savesub", C.c_int32), ("useSubgrid", C.c_int32), ("offset", C.c_double), ("tNStart", C.c_double),
        ("outputFolder", C.c_char * 512),
    ]


def _lib():
    lib = abi.load_library()
    if not getattr(lib, "_case_sigs", False):
        d = C.POINTER(C.c_double)
        sig = {
            "frehd_case_load": ([C.c_char_p, C.c_char_p, C.POINTER(V)], C.c_int),
            "frehd_case_inputs": ([V, C.POINTER(abi.FrehdConfig), C.POINTER(abi.FrehdForcing),
                                   C.POINTER(abi.FrehdIC)], C.c_int),
            "frehd_case_get_info": ([V, C.POINTER(CaseInfo)], C.c_int),
            "frehd_case_write_output": ([V, V, C.c_int32], C.c_int),
            "frehd_case_write_arrays": ([V, C.c_int32] + [d] * 13, C.c_int),
            "frehd_case_free": ([V], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        lib._case_sigs = True
    return lib


def _arr(ptr, n, dtype=np.float64):
    if not ptr or n <= 0:
        return None
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Case:
    """A loaded input deck.  Folders in data.dat are relative to base_dir."""

    def __init__(self, data_dat: str, base_dir: str = ""):
        self.lib = _lib()
        h = V()
        abi.check(self.lib.frehd_case_load(data_dat.encode(), base_dir.encode(), C.byref(h)))
        self.h = h
        self.cfg = abi.FrehdConfig()
        self.forcing = abi.FrehdForcing()
        self.ic = abi.FrehdIC()
        abi.check(self.lib.frehd_case_inputs(self.h, C.byref(self.cfg), C.byref(self.forcing), C.byref(self.ic)))
        self.info = CaseInfo()
        abi.check(self.lib.frehd_case_get_info(self.h, C.byref(self.info)))

    def close(self):
        if getattr(self, "h", None):
            self.lib.frehd_case_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config_dict(self) -> dict:
        return {name: getattr(self.cfg, name) for name, _ in abi.FrehdConfig._fields_}

    def scenario(self) -> cases.Scenario:
        """The loaded inputs as a Scenario (copies), for the oracle and the
        reference harness."""
        k = self.cfg
        N, Nt1 = k.nx * k.ny, k.Nt + 1
        f, ic = self.forcing, self.ic
        zeros = np.zeros(Nt1)
        tideP = _arr(f.tideP, Nt1)
        sc = cases.Scenario(
            cfg=self.config_dict(), bottomZ=_arr(ic.bottomZ, N),
            tideP=tideP if tideP is not None else zeros.copy(),
            tideM=_arr(f.tideM, Nt1) if f.tideM else zeros.copy(),
            inflow=_arr(f.inflow, Nt1) if f.inflow else zeros.copy(),
            tidalPS=_arr(f.tidalPS, Nt1) if f.tidalPS else zeros.copy(),
            tidalMS=_arr(f.tidalMS, Nt1) if f.tidalMS else zeros.copy(),
            tideLocP=_arr(f.tideLocP, k.tideLocLengthP, np.int32) if k.tideLocLengthP else np.zeros(0, np.int32),
            tideLocM=_arr(f.tideLocM, k.tideLocLengthM, np.int32) if k.tideLocLengthM else np.zeros(0, np.int32),
            inflowLoc=_arr(f.inflowLoc, k.inflowLocLength, np.int32) if k.inflowLocLength else np.zeros(0, np.int32),
            manningCell=None, surf0=ic.surf0, U0=ic.U0, V0=ic.V0, S0=ic.S0, S0cell=_arr(ic.S0cell, N),
            name="case", surf0cell=_arr(ic.surf0cell, N), U0cell=_arr(ic.U0cell, N),
            V0cell=_arr(ic.V0cell, N))
        if k.useWind == 1:
            sc.windspd = _arr(f.windspd, Nt1)
            sc.winddir = _arr(f.winddir, Nt1)
        if k.useSubgrid == 1 and ic.sub:
            s = ic.sub.contents
            sg = {"edgeX": _arr(s.edgeX, N), "edgeY": _arr(s.edgeY, N)}
            for t in abi.SUB_TABLES + abi.DRAG_TABLES + abi.MINA_TABLES:
                p = getattr(s, t)
                if p:
                    sg[t] = _arr(p, N * k.Ne)
            for m, src in (("minVxp", "Vxp"), ("minVxm", "Vxm"), ("minVyp", "Vyp"), ("minVym", "Vym")):
                sg.setdefault(m, sg[src])
            sc.subgrid = sg
        return sc

    def write_output(self, dev, tt: int):
        abi.check(self.lib.frehd_case_write_output(self.h, dev.h, int(tt)))

    def write_arrays(self, tt: int, fields: dict):
        d = C.POINTER(C.c_double)
        names = ["surf", "uuXP", "vvYP", "depth", "CDXP", "CDYP", "S", "Nx", "Oy", "Vx", "Vy", "V", "Z"]
        keep = [np.ascontiguousarray(fields[n], dtype=np.float64) if fields.get(n) is not None else None
                for n in names]
        ptrs = [k.ctypes.data_as(d) if k is not None else None for k in keep]
        abi.check(self.lib.frehd_case_write_arrays(self.h, int(tt), *ptrs))


def run_case(data_dat: str, base_dir: str = "", steps: int | None = None, precision=abi.PREC_FP64,
             warm_start: int = 0, log=print):
    """The reference's time loop on the device: writes the initial state
    (step 0) and every OutItvl steps; returns the handle after `steps` steps
    (default: the deck's Nt)."""
    from .device import DeviceScenario
    case = Case(data_dat, base_dir)
    sc = case.scenario()
    dev = DeviceScenario(sc, precision=precision, warmStart=warm_start)
    dev._case = case
    total = case.info.Nt if steps is None else min(int(steps), case.info.Nt)
    every = case.info.OutItvl
    case.write_output(dev, 0)
    done = 0
    while done < total:
        n = min(every - done % every, total - done)
        rc = dev.try_step(n)
        done += n
        if rc == abi.FREHD_ECFL:
            case.write_output(dev, done)
            log(f"WARNING: Max CFL exceeds 20, stopped at step {dev.tt}")
            break
        abi.check(rc)
        if done % every == 0:
            case.write_output(dev, done)
    return dev
