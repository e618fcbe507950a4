"""Subgrid bathymetry tables from a fine bathymetry.

Restates the reference's table generator (preProcessing/
script_CreateSubAreaWithBlockV16.m:129-262 with checkBlock = 0, face areas as
in ComputeSubgridFaceArea.m:1-43) for a fine grid of r x r pixels per coarse
cell.  Orientation: MATLAB's first index is the C x index (the reference
reshapes column-major); here arrays are [y, x], so "p" halves and faces are
the east (x) and north (y) sides.

Level k is surfmin + k*dsurf for k = 0 .. Ne-1 with
Ne = round((surfmax - surfmin)/dsurf) (subgrid.c:2116: the reference reads Ne
levels of the Ne+1 the MATLAB script writes).  Tables are level-major,
t[k*N + ii] with ii = j*nx + i (subgrid.c:2100-2107).
"""
from __future__ import annotations

import numpy as np


def half_split(r: int) -> int:
    """MATLAB round(r/2) (half away from zero)."""
    return (r + 1) // 2


def coarse_geometry(fine: np.ndarray, r: int):
    """bottom = min over the block; edgeX/edgeY = face bottoms: the higher of
    the two pixel columns' (rows') minima straddling the face
    (script V16 :129-170)."""
    fy, fx = fine.shape
    ny, nx = fy // r, fx // r
    b = fine.reshape(ny, r, nx, r)
    bottom = b.min(axis=(1, 3))
    # east face: fine columns x = i*r + r - 1 and i*r + r, over the block's rows
    colmin = fine.reshape(ny, r, fx).min(axis=1)           # (ny, fx): min over rows of each column
    last = colmin[:, r - 1::r]                              # (ny, nx)
    nxt = np.empty_like(last)
    nxt[:, :-1] = colmin[:, r::r]
    edgeX = np.maximum(last, nxt)
    edgeX[:, -1] = last[:, -1]
    rowmin = fine.reshape(fy, nx, r).min(axis=2)           # (fy, nx)
    lastr = rowmin[r - 1::r, :]
    nxtr = np.empty_like(lastr)
    nxtr[:-1, :] = rowmin[r::r, :]
    edgeY = np.maximum(lastr, nxtr)
    edgeY[-1, :] = lastr[-1, :]
    return bottom, edgeX, edgeY


def tables(fine: np.ndarray, r: int, dxf: float, surfmin: float, dsurf: float, Ne: int,
           dtype=np.float64):
    """Level tables V, Z, Vxp, Vxm, Vyp, Vym, Np, Nm, Op, Om (and minV*, set to
    the half volumes, their value when no minimum-area data is used)."""
    fy, fx = fine.shape
    ny, nx = fy // r, fx // r
    N = nx * ny
    dA = dxf * dxf
    h2 = half_split(r)
    out = {k: np.zeros((Ne, N), dtype=dtype) for k in
           ("V", "Z", "Vxp", "Vxm", "Vyp", "Vym", "Np", "Nm", "Op", "Om")}
    for lev in range(Ne):
        h = surfmin + lev * dsurf
        D = np.maximum(h - fine, 0.0)
        B = D.reshape(ny, r, nx, r)
        out["V"][lev] = (dA * B.sum(axis=(1, 3))).reshape(N)
        out["Z"][lev] = (dA * (B > 0).sum(axis=(1, 3))).reshape(N)
        if r > 1:
            out["Vxp"][lev] = (dA * B[:, :, :, h2:].sum(axis=(1, 3))).reshape(N)
            out["Vxm"][lev] = (dA * B[:, :, :, :h2].sum(axis=(1, 3))).reshape(N)
            out["Vyp"][lev] = (dA * B[:, h2:, :, :].sum(axis=(1, 3))).reshape(N)
            out["Vym"][lev] = (dA * B[:, :h2, :, :].sum(axis=(1, 3))).reshape(N)
            # faces: sum over the face's pixels of the shallower side's depth
            a = D[:, r - 1::r]                                   # (fy, nx) last column of each cell
            bnext = np.empty_like(a)
            bnext[:, :-1] = D[:, r::r]
            m = np.minimum(a, bnext)
            m[:, -1] = a[:, -1]
            Np = m.reshape(ny, r, nx).sum(axis=1) * dxf
            Nm = np.empty_like(Np)
            Nm[:, 1:] = Np[:, :-1]
            Nm[:, 0] = D[:, 0].reshape(ny, r).sum(axis=1) * dxf
            a = D[r - 1::r, :]                                   # (ny, fx) last row of each cell
            bnext = np.empty_like(a)
            bnext[:-1, :] = D[r::r, :]
            m = np.minimum(a, bnext)
            m[-1, :] = a[-1, :]
            Op = m.reshape(ny, nx, r).sum(axis=2) * dxf
            Om = np.empty_like(Op)
            Om[1:, :] = Op[:-1, :]
            Om[0, :] = D[0, :].reshape(nx, r).sum(axis=1) * dxf
            out["Np"][lev], out["Nm"][lev] = Np.reshape(N), Nm.reshape(N)
            out["Op"][lev], out["Om"][lev] = Op.reshape(N), Om.reshape(N)
        else:
            d = D.reshape(N)
            for k in ("Vxp", "Vxm", "Vyp", "Vym"):
                out[k][lev] = dA * d / 2
            for k in ("Np", "Nm", "Op", "Om"):
                out[k][lev] = dxf * d
    for k, src in (("minVxp", "Vxp"), ("minVxm", "Vxm"), ("minVyp", "Vyp"), ("minVym", "Vym")):
        out[k] = out[src]
    return {k: np.ascontiguousarray(v.reshape(-1)) for k, v in out.items()}


def _spec(fine, r, dxf, surfmin, dsurf, Ne, check_block, min_volume, yh, cn, n0):
    from . import abi
    fy, fx = fine.shape
    return abi.FrehdSubgridSpec(nxFine=fx, nyFine=fy, r=r, Ne=Ne, dxf=dxf, surfmin=surfmin, dsurf=dsurf,
                                checkBlock=int(check_block), useminVolume=int(min_volume), useYh=int(yh),
                                useCn=int(cn), n0=n0)


def generate(fine: np.ndarray, r: int, dxf: float, surfmin: float, dsurf: float, Ne: int, check_block=True,
             min_volume=True, yh=False, cn=False, n0=0.03, threads=0):
    """The reference's table generator with its block check (native host
    code, frehd_subgrid_generate; preProcessing/script_CreateSubAreaWithBlockV16.m
    with ComputeFaceAreaWithBlockCheck.m): level tables (level-major, N =
    nx * ny per level) and the cell / face bottoms.  Returns a dict of flat
    float64 arrays: the ten tables, Nmin Nmax Omin Omax (check_block),
    minVxp minVxm minVyp minVym (min_volume), YhX YhY (yh), Cn (cn), bottom
    bottomXP bottomYP.  `fine` is [y, x]."""
    from . import abi
    import ctypes as C
    lib = abi.load_library()
    fine = np.ascontiguousarray(fine, dtype=np.float64)
    fy, fx = fine.shape
    N = (fy // r) * (fx // r)
    want = list(abi.SUBGRID_TABLES[:10])
    if check_block:
        want += ["Nmin", "Nmax", "Omin", "Omax"]
    if min_volume:
        want += ["minVxp", "minVxm", "minVyp", "minVym"]
    if yh:
        want += ["YhX", "YhY"]
    if cn:
        want += ["Cn"]
    arrays = {k: np.zeros(N * Ne) for k in want}
    arrays.update({k: np.zeros(N) for k in abi.SUBGRID_GEOMETRY})
    out = abi.FrehdSubgridOut(**{k: v.ctypes.data_as(abi.c_double_p) for k, v in arrays.items()})
    spec = _spec(fine, r, dxf, surfmin, dsurf, Ne, check_block, min_volume, yh, cn, n0)
    abi.check(lib.frehd_subgrid_generate(fine.ctypes.data_as(abi.c_double_p), C.byref(spec), C.byref(out),
                                         int(threads)))
    return arrays


def write_text(folder: str, tabs: dict, fine_shape, r: int, dxf: float, surfmin: float, dsurf: float, Ne: int):
    """Write `tabs` (generate()'s dict) as the text files initAllSubVar reads
    (frehd_subgrid_write_text: "<folder>/<name>.dat", one "%f" line per
    value)."""
    from . import abi
    import ctypes as C
    lib = abi.load_library()
    fy, fx = fine_shape
    spec = abi.FrehdSubgridSpec(nxFine=fx, nyFine=fy, r=r, Ne=Ne, dxf=dxf, surfmin=surfmin, dsurf=dsurf)
    keep = {k: np.ascontiguousarray(v, dtype=np.float64) for k, v in tabs.items()
            if k in abi.SUBGRID_TABLES + abi.SUBGRID_GEOMETRY}
    out = abi.FrehdSubgridOut(**{k: v.ctypes.data_as(abi.c_double_p) for k, v in keep.items()})
    abi.check(lib.frehd_subgrid_write_text(folder.encode(), C.byref(spec), C.byref(out)))
