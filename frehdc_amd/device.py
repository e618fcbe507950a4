"""Python handle over the device library (libfrehd_gpu.so).

Mirrors the reference's time loop (FrehdC_v42/nssolve.c:48-106): create a
handle from a scenario (Init), advance steps (oneCompleteStep + monitorCFL),
download the state (DataOutput's arrays at full precision).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


class DeviceScenario:
    """One scenario resident on the GPU."""

    def __init__(self, scen, from_state: dict | None = None, **override):
        self.lib = abi.load_library()
        self.scen = scen
        self._cfg = scen.c_config(**override)
        self._forc = scen.c_forcing()
        self.nScalar = max(1, self._cfg.nScalar)
        h = C.c_void_p()
        if from_state is None:
            self._ic = scen.c_ic()
            rc = self.lib.frehd_dev_create(C.byref(self._cfg), C.byref(self._forc),
                                           C.byref(self._ic), C.byref(h))
        else:
            keep = {k: np.ascontiguousarray(v, dtype=np.int32 if k == "ind" else np.float64)
                    for k, v in from_state.items()}
            self._keep = keep
            flds = abi.FrehdFields()
            for name in abi.FIELD_NAMES + abi.SUB_FIELD_NAMES:
                if name in keep:
                    setattr(flds, name, abi.dptr(keep[name]))
            if "ind" in keep:
                flds.ind = abi.iptr(keep["ind"])
            self._ic = scen.c_ic()
            sub = self._ic.sub if scen.subgrid is not None else None
            rc = self.lib.frehd_dev_create_from_state(C.byref(self._cfg), C.byref(self._forc),
                                                      sub, C.byref(flds), C.byref(h))
        abi.check(rc)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.frehd_dev_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1) -> int:
        """Advance n steps; returns 0 or raises FrehdError (FREHD_ECFL when
        the CFL stop test fires)."""
        return abi.check(self.lib.frehd_dev_step(self.h, int(n)))

    def try_step(self, n: int = 1) -> int:
        return self.lib.frehd_dev_step(self.h, int(n))

    def phase(self, p: int) -> int:
        return abi.check(self.lib.frehd_dev_phase(self.h, int(p)))

    def download(self, names=None) -> dict:
        s = self.scen
        names = list(abi.FIELD_NAMES) if names is None else list(names)
        out = {}
        flds = abi.FrehdFields()
        for n in names:
            if n == "ind":
                a = np.zeros(s.Ng, dtype=np.int32)
                out[n] = a
                flds.ind = abi.iptr(a)
                continue
            if n == "S":
                a = np.zeros(self.nScalar * s.Ng)
            elif n == "Sm":
                a = np.zeros(self.nScalar * s.N)
            elif n in abi.INTERIOR_FIELDS:
                a = np.zeros(s.N)
            else:
                a = np.zeros(s.Ng)
            out[n] = a
            setattr(flds, n, abi.dptr(a))
        abi.check(self.lib.frehd_dev_download(self.h, C.byref(flds)))
        if "S" in out:
            out["S"] = out["S"].reshape(self.nScalar, s.Ng)
        if "Sm" in out:
            out["Sm"] = out["Sm"].reshape(self.nScalar, s.N)
        return out

    def field(self, name):
        return self.download([name])[name]

    def upload(self, fields: dict):
        """frehd_dev_upload: replace the named state arrays (the layout of
        download(); S as nScalar x Ng) at a step boundary."""
        flds = abi.FrehdFields()
        keep = []
        for n, a in fields.items():
            if n == "ind":
                a = np.ascontiguousarray(a, dtype=np.int32).reshape(-1)
                flds.ind = abi.iptr(a)
            else:
                a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                setattr(flds, n, abi.dptr(a))
            keep.append(a)
        abi.check(self.lib.frehd_dev_upload(self.h, C.byref(flds)))

    def stats(self) -> dict:
        st = abi.FrehdStats()
        abi.check(self.lib.frehd_dev_stats(self.h, C.byref(st)))
        return dict(steps=st.steps, pcgIterTotal=st.pcgIterTotal, pcgIterLast=st.pcgIterLast,
                    refineLast=st.refineLast, relResLast=st.relResLast, maxCFLx=st.maxCFLx,
                    maxCFLy=st.maxCFLy, Vloss=list(st.Vloss), halted=bool(st.halted),
                    haltStep=st.haltStep)

    def set_budget(self, on=True):
        """Switch the conservation budgets on (cleared) or off
        (frehd_dev_set_budget)."""
        abi.check(self.lib.frehd_dev_set_budget(self.h, 1 if on else 0))

    def budget(self) -> dict:
        """The conservation budgets (frehd_dev_budget): volume and, per
        scalar, mass residuals of the flux-form updates, and relative
        measures (largest step residual over the total, summed residual
        over the total)."""
        b = abi.FrehdBudget()
        abi.check(self.lib.frehd_dev_budget(self.h, C.byref(b)))
        ns = b.nScalar
        out = dict(steps=b.steps, on=bool(b.on), volume=b.volume, volumeResidual=b.volumeResidual,
                   volumeResidualMax=b.volumeResidualMax, volumeBoundary=b.volumeBoundary,
                   volumeInflow=b.volumeInflow, volumeClip=b.volumeClip, volumeMoved=b.volumeMoved,
                   mass=list(b.mass[:ns]), massResidual=list(b.massResidual[:ns]),
                   massResidualMax=list(b.massResidualMax[:ns]), massBoundary=list(b.massBoundary[:ns]),
                   massInflow=list(b.massInflow[:ns]), massMoved=list(b.massMoved[:ns]))
        out["volumeRelResidual"] = abs(b.volumeResidual) / b.volume if b.volume > 0 else 0.0
        out["volumeRelResidualMax"] = b.volumeResidualMax / b.volume if b.volume > 0 else 0.0
        # a scalar's mass may start at zero (the boundary brings it in): its
        # scale is the larger of the total and the mass the faces moved
        scale = [max(abs(m), abs(mv) / max(b.steps, 1), 1e-300) for m, mv in
                 zip(out["mass"], out["massMoved"])]
        out["massRelResidual"] = [abs(r) / s for r, s in zip(out["massResidual"], scale)]
        out["massRelResidualMax"] = [r / s for r, s in zip(out["massResidualMax"], scale)]
        return out

    @property
    def tt(self):
        v = C.c_int32()
        abi.check(self.lib.frehd_dev_step_index(self.h, C.byref(v)))
        return v.value

    @tt.setter
    def tt(self, value):
        abi.check(self.lib.frehd_dev_set_step_index(self.h, int(value)))

    def tables(self) -> dict:
        """The handle's subgrid level tables, {name: (Ne, N) array}."""
        cfg = self._cfg
        N = cfg.nx * cfg.ny
        out = {k: np.empty((cfg.Ne, N)) for k in abi.SUB_TABLES[:10]}
        ptrs = (abi.c_double_p * 10)(*[abi.dptr(out[k]) for k in abi.SUB_TABLES[:10]])
        abi.check(self.lib.frehd_dev_download_tables(self.h, ptrs))
        return out

    def set_pcg_tiling(self, rows_per_tile=0, xcd_remap=False, cols_per_lane=1):
        abi.check(self.lib.frehd_dev_set_pcg_tiling(self.h, int(rows_per_tile), 1 if xcd_remap else 0,
                                                    int(cols_per_lane)))

    def autotune_blocks(self, steps: int = 2) -> dict:
        """frehd_dev_autotune_blocks: the stencil kernels' block shape by
        timing; the model advances by 6 * steps steps."""
        return _autotune_blocks(self.lib.frehd_dev_autotune_blocks, self.h, steps)

    def set_fused_solve(self, max_blocks: int = 0):
        """Mixed form: the whole solve as one persistent launch (0: as many
        blocks as the device holds, > 0: at most that many, < 0: off, the
        launched solve) -- frehd_dev_set_fused_solve."""
        abi.check(self.lib.frehd_dev_set_fused_solve(self.h, int(max_blocks)))

    def autotune(self, iters: int = 20) -> dict:
        """Time the solver under every candidate tiling on the last step's
        system and keep the fastest (frehd_dev_autotune); the state is left
        as it was."""
        return _autotune(self.lib.frehd_dev_autotune, self.h, iters)

    def enable_timing(self, on=True):
        abi.check(self.lib.frehd_dev_enable_timing(self.h, 1 if on else 0))

    def kernel_times(self) -> dict:
        cap = 64
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        launches = (C.c_int64 * cap)()
        n = C.c_int32()
        abi.check(self.lib.frehd_dev_kernel_times(self.h, cap, names, ms, launches, C.byref(n)))
        return {names[q].decode(): (ms[q], launches[q]) for q in range(min(n.value, cap))}


class Batch:
    """Many DeviceScenario handles advanced by shared kernel launches
    (frehd_batch_*): members may differ in grid size, forcing and physics but
    share useSubgrid and precision."""

    def __init__(self, capacity: int):
        self.lib = abi.load_library()
        h = C.c_void_p()
        abi.check(self.lib.frehd_batch_create(int(capacity), C.byref(h)))
        self.h = h
        self.members = []

    def add(self, dev: DeviceScenario):
        abi.check(self.lib.frehd_batch_add(self.h, dev.h))
        self.members.append(dev)

    def remove(self, dev: DeviceScenario):
        abi.check(self.lib.frehd_batch_remove(self.h, dev.h))
        self.members.remove(dev)

    def __len__(self):
        n = C.c_int32()
        abi.check(self.lib.frehd_batch_size(self.h, C.byref(n)))
        return n.value

    def step(self, n: int = 1) -> int:
        """Advance every member n steps; returns 0 or raises FrehdError
        (FREHD_ECFL when a member is halted)."""
        return abi.check(self.lib.frehd_batch_step(self.h, int(n)))

    def try_step(self, n: int = 1) -> int:
        return self.lib.frehd_batch_step(self.h, int(n))

    def autotune_blocks(self, steps: int = 2) -> dict:
        """frehd_batch_autotune_blocks: one stencil block shape for all
        members; they advance by 6 * steps steps."""
        return _autotune_blocks(self.lib.frehd_batch_autotune_blocks, self.h, steps)

    def set_fused_solve(self, max_blocks: int = 0):
        """frehd_batch_set_fused_solve: one fused solve launch for all members."""
        abi.check(self.lib.frehd_batch_set_fused_solve(self.h, int(max_blocks)))

    def autotune(self, iters: int = 20) -> dict:
        """frehd_batch_autotune: one solver tiling for all members."""
        return _autotune(self.lib.frehd_batch_autotune, self.h, iters)

    def enable_timing(self, on=True):
        abi.check(self.lib.frehd_batch_enable_timing(self.h, 1 if on else 0))

    def kernel_times(self) -> dict:
        cap = 64
        names = (C.c_char_p * cap)()
        ms = (C.c_double * cap)()
        launches = (C.c_int64 * cap)()
        n = C.c_int32()
        abi.check(self.lib.frehd_batch_kernel_times(self.h, cap, names, ms, launches, C.byref(n)))
        return {names[q].decode(): (ms[q], launches[q]) for q in range(min(n.value, cap))}

    def close(self):
        if getattr(self, "h", None):
            self.lib.frehd_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _autotune_blocks(fn, h, steps):
    import ctypes as C
    bx, ms = C.c_int32(), C.c_double()
    abi.check(fn(h, int(steps), C.byref(bx), C.byref(ms)))
    return {"block_x": bx.value, "block_y": 256 // bx.value, "ms_per_step": ms.value}


def fused_solve_capacity(cols_per_lane: int = 4) -> int:
    """Blocks of the fused solve the device holds at once."""
    import ctypes as C
    n = C.c_int32()
    abi.check(abi.load_library().frehd_fused_solve_capacity(int(cols_per_lane), C.byref(n)))
    return n.value


def device_count() -> int:
    lib = abi.load_library()
    n = C.c_int32()
    lib.frehd_device_count(C.byref(n))
    return n.value


def _autotune(fn, h, iters):
    H, remap, lanes, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
    abi.check(fn(h, int(iters), C.byref(H), C.byref(remap), C.byref(lanes), C.byref(ms)))
    return {"rows_per_tile": H.value, "xcd_remap": remap.value, "cols_per_lane": lanes.value,
            "ms_per_iteration": ms.value}
