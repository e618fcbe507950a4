"""CPU restatement of the reference's subgrid table generator, for small
beds.  TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the native
generator (frehd_subgrid_generate); the product never imports it.

Follows preProcessing/script_CreateSubAreaWithBlockV16.m:144-460 (cell and
face bottoms, volumes, wet area, half volumes, blocked faces zeroed, Yh, Cn,
minimum volumes) with ComputeFaceAreaWithBlockCheck.m:1-326 (checkBlock = 1)
or ComputeSubgridFaceArea.m:1-38 (checkBlock = 0), statement by statement,
on arrays in the script's orientation: B[X, Y] with X the first index (the
reference's x), 0-based.  Connected wet clusters come from
scipy.ndimage.label on the transposed cell: its row-major numbering of the
transpose is the column-major numbering bwlabel uses (clusters numbered in
the order of their first pixel scanning the first index fastest).

Sums follow the script: sum(M(:)) of a cell as one column-major vector,
sum(sum(M)) and sum(M, d) over the first index of each column first.  The
tests compare on beds and levels that are multiples of a power of two,
where every sum is exact and its order immaterial, so the comparison is
exact.

Parity: pinned by closed forms (flat beds), by the plain generator
(frehdc_amd/subgrid.py tables(), itself pinned by brute force) where nothing
is blocked, and by constructed beds whose blocking, cleaning and minimum
areas are worked out by hand (tests/test_subgrid_generator.py).  MATLAB is
absent, so the script's own outputs are not available as vectors; bwlabel's
numbering (which decides only between equally large clusters) is taken from
its documentation.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import ndimage


def _colsum(M):
    """sum(M, 1): over the first index, per column (sequential)."""
    out = []
    for b in range(M.shape[1]):
        s = 0.0
        for a in range(M.shape[0]):
            s += float(M[a, b])
        out.append(s)
    return out


def _rowsum(M):
    """sum(M, 2): over the second index, per row (sequential)."""
    out = []
    for a in range(M.shape[0]):
        s = 0.0
        for b in range(M.shape[1]):
            s += float(M[a, b])
        out.append(s)
    return out


def _vecsum(M):
    """sum(M(:)): column-major."""
    s = 0.0
    for b in range(M.shape[1]):
        for a in range(M.shape[0]):
            s += float(M[a, b])
    return s


def _sumsum(M):
    """sum(sum(M))"""
    s = 0.0
    for v in _colsum(M):
        s += v
    return s


def _median(v):
    w = sorted(v)
    n = len(w)
    return w[n // 2] if n % 2 else (w[n // 2 - 1] + w[n // 2]) / 2


def _mround(x):
    return int(math.floor(x + 0.5))


def _block_check(B, D, h, r, dx, Dim, tabs, k):
    """ComputeFaceAreaWithBlockCheck.m for level k: returns the block flags
    and writes Np Nm Op Om Nmin Nmax Omin Omax into tabs[name][X, Y, k]."""
    d1, d2 = Dim
    dim = B.shape
    blk = {q: np.zeros(Dim, dtype=bool) for q in ("Np", "Nm", "Op", "Om")}
    binary = (h > B).astype(np.int8)
    # steps 1 and 2 (:28-77)
    for ii in range(d1):
        for jj in range(d2):
            cell = binary[ii * r:(ii + 1) * r, jj * r:(jj + 1) * r].copy()
            lab_t, nw = ndimage.label(cell.T)          # 4-connectivity
            lab = lab_t.T
            if nw > 0:
                sizes = [int((lab == p).sum()) for p in range(1, nw + 1)]
                ind = int(np.argmax(sizes)) + 1       # first of the largest
                bb = lab == ind
                if bb.sum() < 0.01 * r * r:
                    cell[:] = 0
                else:
                    cell[bb] = 1
                    cell[~bb] = 0
            else:
                cell[:] = 0
            if not cell[0, :].any():
                blk["Nm"][ii, jj] = True
            if not cell[-1, :].any():
                blk["Np"][ii, jj] = True
            if not cell[:, 0].any():
                blk["Om"][ii, jj] = True
            if not cell[:, -1].any():
                blk["Op"][ii, jj] = True
            binary[ii * r:(ii + 1) * r, jj * r:(jj + 1) * r] = cell
    # step 3 (:80-118), from the second row and column
    for ii in range(1, d1):
        for jj in range(1, d2):
            this = binary[ii * r:(ii + 1) * r, jj * r:(jj + 1) * r]
            cNm = binary[(ii - 1) * r:ii * r, jj * r:(jj + 1) * r]
            cOm = binary[ii * r:(ii + 1) * r, (jj - 1) * r:jj * r]
            if not any(this[0, p] == 1 and cNm[-1, p] == 1 for p in range(r)):
                blk["Nm"][ii, jj] = True
                blk["Np"][ii - 1, jj] = True
            if not any(this[p, 0] == 1 and cOm[p, -1] == 1 for p in range(r)):
                blk["Om"][ii, jj] = True
                blk["Op"][ii, jj - 1] = True
    # step 4 (:120-166), masks from the cleaned cells
    mask = binary
    Np, Nm, Op, Om = (tabs[q] for q in ("Np", "Nm", "Op", "Om"))
    for ii in range(d1):
        for jj in range(d2):
            i1, i2 = ii * r, ii * r + r - 1
            j1, j2 = jj * r, jj * r + r - 1
            if i2 != dim[0] - 1:
                dNp = np.where(mask[i2, j1:j2 + 1] == 0, 0.0, D[i2, j1:j2 + 1])
                dNm = np.where(mask[i2 + 1, j1:j2 + 1] == 0, 0.0, D[i2 + 1, j1:j2 + 1])
                s = 0.0
                for v in np.minimum(dNp, dNm):
                    s += float(v)
                Np[ii, jj, k] = s * dx
                Nm[ii + 1, jj, k] = s * dx
            else:
                Np[ii, jj, k] = sum(float(v) for v in D[i2, j1:j2 + 1]) * dx
                Nm[0, jj, k] = sum(float(v) for v in D[0, j1:j2 + 1]) * dx
            if j2 != dim[1] - 1:
                dOp = np.where(mask[i1:i2 + 1, j2] == 0, 0.0, D[i1:i2 + 1, j2])
                dOm = np.where(mask[i1:i2 + 1, j2 + 1] == 0, 0.0, D[i1:i2 + 1, j2 + 1])
                s = 0.0
                for v in np.minimum(dOp, dOm):
                    s += float(v)
                Op[ii, jj, k] = s * dx
                Om[ii, jj + 1, k] = s * dx
            else:
                Op[ii, jj, k] = sum(float(v) for v in D[i1:i2 + 1, j2]) * dx
                Om[ii, 0, k] = sum(float(v) for v in D[i1:i2 + 1, 0]) * dx
    # step 5 (:169-322): minimum and maximum areas
    fl, rd = r // 2, _mround(r / 2)
    fl3, rd3 = (3 * r) // 4, _mround(3 * r / 4)
    nex_prev = oex_prev = None
    for ii in range(d1):
        for jj in range(d2):
            i1, i2 = ii * r, ii * r + r - 1
            j1, j2 = jj * r, jj * r + r - 1
            # x: no mask (commented out in the script)
            if i2 != dim[0] - 1:
                dN = D[i2 - fl + 1:i2 + rd + 1, j1:j2 + 1]
                dNex = D[i2 - fl3 + 1:i2 + rd3 + 1, j1:j2 + 1]
                nex_prev = dNex
            else:
                dN = D[i2 - fl + 1:i2 + 1, j1:j2 + 1]
                dNex = nex_prev                          # not reassigned there
            sN = _rowsum(dN)
            sNex = _rowsum(dNex) if dNex is not None else [0.0]
            bb = [v == 0 for v in sN]
            kept = [v for v in sN if v != 0]
            if kept:
                nmin, nminex, nmax, nmed = min(kept), min(sNex), max(kept), _median(kept)
                nmid = Np[ii, jj, k]
                tabs["Nmin"][ii, jj, k] = nmin
                cc = [v == nminex for v in sNex]
                if (i2 != dim[0] - 1 and not bb[0] and not bb[-1] and Np[ii + 1, jj, k] > 0 and Nm[ii, jj, k] > 0
                        and len(kept) > 0.5 * r and not cc[0] and not cc[-1]):
                    if len(kept) > 3:
                        if abs(nmax - nmed) > 2 * abs(nmed - nmin):
                            tabs["Nmax"][ii, jj, k] = nmin
                        elif abs(nmed - nmin) > 2 * abs(nmax - nmed):
                            tabs["Nmax"][ii, jj, k] = nmin
                        else:
                            tabs["Nmax"][ii, jj, k] = nmid
                    else:
                        tabs["Nmax"][ii, jj, k] = nmid
                else:
                    tabs["Nmax"][ii, jj, k] = nmid
            # y: the cleaned-away pixels masked
            if j2 != dim[1] - 1:
                dO = D[i1:i2 + 1, j2 - fl + 1:j2 + rd + 1].copy()
                mO = mask[i1:i2 + 1, j2 - fl + 1:j2 + rd + 1]
                dOex = D[i1:i2 + 1, j2 - fl3 + 1:j2 + rd3 + 1]
                oex_prev = dOex
            else:
                dO = D[i1:i2 + 1, j2 - fl + 1:j2 + 1].copy()
                mO = mask[i1:i2 + 1, j2 - fl + 1:j2 + 1]
                dOex = oex_prev
            dO[mO == 0] = 0.0
            sO = _colsum(dO)
            sOex = _colsum(dOex) if dOex is not None else [0.0]
            bb = [v == 0 for v in sO]
            kept = [v for v in sO if v != 0]
            if kept:
                omin, ominex, omax, omed = min(kept), min(sOex), max(kept), _median(kept)
                omid = Op[ii, jj, k]
                tabs["Omin"][ii, jj, k] = omin
                cc = [v == ominex for v in sOex]
                if (j2 != dim[1] - 1 and not bb[0] and not bb[-1] and Op[ii, jj + 1, k] > 0 and Om[ii, jj, k] > 0
                        and len(kept) > 0.5 * r and not cc[0] and not cc[-1]):
                    if len(kept) > 3:
                        if abs(omax - omed) > 2 * abs(omed - omin):
                            tabs["Omax"][ii, jj, k] = omin
                        elif abs(omed - omin) > 2 * abs(omax - omed):
                            tabs["Omax"][ii, jj, k] = omin
                        else:
                            tabs["Omax"][ii, jj, k] = omid
                    else:
                        tabs["Omax"][ii, jj, k] = omid
                else:
                    tabs["Omax"][ii, jj, k] = omid
    return blk


def _plain_faces(D, r, dx, Dim, tabs, k):
    """ComputeSubgridFaceArea.m:8-38"""
    d1, d2 = Dim
    dim = D.shape
    for ii in range(d1):
        for jj in range(d2):
            i1, i2 = ii * r, ii * r + r - 1
            j1, j2 = jj * r, jj * r + r - 1
            if i2 != dim[0] - 1:
                s = 0.0
                for v in np.minimum(D[i2, j1:j2 + 1], D[i2 + 1, j1:j2 + 1]):
                    s += float(v)
                tabs["Np"][ii, jj, k] = s * dx
                tabs["Nm"][ii + 1, jj, k] = s * dx
            else:
                tabs["Np"][ii, jj, k] = sum(float(v) for v in D[i2, j1:j2 + 1]) * dx
                tabs["Nm"][0, jj, k] = sum(float(v) for v in D[0, j1:j2 + 1]) * dx
            if j2 != dim[1] - 1:
                s = 0.0
                for v in np.minimum(D[i1:i2 + 1, j2], D[i1:i2 + 1, j2 + 1]):
                    s += float(v)
                tabs["Op"][ii, jj, k] = s * dx
                tabs["Om"][ii, jj + 1, k] = s * dx
            else:
                tabs["Op"][ii, jj, k] = sum(float(v) for v in D[i1:i2 + 1, j2]) * dx
                tabs["Om"][ii, 0, k] = sum(float(v) for v in D[i1:i2 + 1, 0]) * dx


def generate(fine, r, dx, surfmin, dsurf, Ne, check_block=True, min_volume=True, yh=False, cn=False, n0=0.03):
    """The tables of levels surfmin + k dsurf, k < Ne, from `fine` [y, x];
    returned flat, level-major with cell index j*nx + i, like the native
    generator's."""
    B = np.asarray(fine, dtype=np.float64).T          # bathCenter(X, Y)
    dim = B.shape
    Dim = (dim[0] // r, dim[1] // r)
    d1, d2 = Dim
    Dx = r * dx
    dA = dx * dx
    names = ["V", "Z", "Vxp", "Vyp", "Vxm", "Vym", "Np", "Op", "Nm", "Om", "Nmin", "Omin", "Nmax", "Omax",
             "minVxp", "minVxm", "minVyp", "minVym", "YhX", "YhY", "Cn"]
    tabs = {q: np.zeros((d1, d2, Ne)) for q in names}
    bottom = np.zeros(Dim)
    bXP = np.zeros(Dim)
    bYP = np.zeros(Dim)
    # cell bathymetry (:144-172)
    for ii in range(d1):
        for jj in range(d2):
            i1, i2 = ii * r, ii * r + r - 1
            j1, j2 = jj * r, jj * r + r - 1
            bath = B[i1:i2 + 1, j1:j2 + 1]
            bottom[ii, jj] = bath.min()
            lbx = B[i2:i2 + 1, j1:j2 + 1] if i2 == dim[0] - 1 else B[i2:i2 + 2, j1:j2 + 1]
            lby = B[i1:i2 + 1, j2:j2 + 1] if j2 == dim[1] - 1 else B[i1:i2 + 1, j2:j2 + 2]
            bXP[ii, jj] = lbx.min(axis=1).max()
            bYP[ii, jj] = lby.min(axis=0).max()
    h2 = _mround(r / 2)
    for k in range(Ne):
        h = surfmin + k * dsurf
        D = np.maximum(h - B, 0.0)
        blk = None
        if check_block and r > 1:
            blk = _block_check(B, D, h, r, dx, Dim, tabs, k)
        elif r > 1:
            _plain_faces(D, r, dx, Dim, tabs, k)
        for ii in range(d1):
            for jj in range(d2):
                i1, i2 = ii * r, ii * r + r - 1
                j1, j2 = jj * r, jj * r + r - 1
                L = D[i1:i2 + 1, j1:j2 + 1]
                tabs["V"][ii, jj, k] = dA * _vecsum(L)
                tabs["Z"][ii, jj, k] = dA * float((L > 0).sum())
                if r > 1:
                    tabs["Vxp"][ii, jj, k] = dA * _sumsum(L[h2:r, :])
                    tabs["Vyp"][ii, jj, k] = dA * _sumsum(L[:, h2:r])
                    tabs["Vxm"][ii, jj, k] = dA * _sumsum(L[0:h2, :])
                    tabs["Vym"][ii, jj, k] = dA * _sumsum(L[:, 0:h2])
                    if blk is not None:
                        for q in ("Np", "Nm", "Op", "Om"):
                            if blk[q][ii, jj] and tabs[q][ii, jj, k] != 0:
                                tabs[q][ii, jj, k] = 0
                else:
                    for q in ("Vxp", "Vyp", "Vxm", "Vym"):
                        tabs[q][ii, jj, k] = dA * float(L[0, 0]) / 2
                    for q in ("Np", "Op", "Nm", "Om"):
                        tabs[q][ii, jj, k] = dx * float(L[0, 0])
                if yh:
                    Y = tabs["V"][ii, jj, k] / (Dx ** 2)
                    wet = [float(L[a, b]) for b in range(r) for a in range(r) if L[a, b] > 0]
                    if wet:
                        s = 0.0
                        for v in wet:
                            s += v ** (5 / 3)
                        sumh = s ** (-2)
                        tabs["YhX"][ii, jj, k] = (Dx ** 2 / dx ** 2) * (Y ** 3) * sumh
                        tabs["YhY"][ii, jj, k] = (Dx ** 2 / dx ** 2) * (Y ** 3) * sumh
                if cn:
                    cellz = B[i1:i2 + 1, j1:j2 + 1].copy()
                    cellz[cellz > h] = bottom[ii, jj]
                    colmeans = [v / r for v in _colsum(cellz - bottom[ii, jj])]
                    m = 0.0
                    for v in colmeans:
                        m += v
                    Dm = max(0.0, m / r)
                    H = max(0.0, h - bottom[ii, jj])
                    xi = H / Dm if Dm > 0 else 100.0
                    if 0.2 <= xi <= 7:
                        fxi = 1 + (1 / xi) * math.log(math.cosh(1 - xi) / math.cosh(1))
                        tabs["Cn"][ii, jj, k] = H ** (1 / 6) / math.sqrt(9.81 * 4.5 * fxi)
                    else:
                        tabs["Cn"][ii, jj, k] = n0
        if min_volume:
            _min_volume(D, r, dx, Dx, Dim, tabs, k, h2)
    out = {}
    for q in names:
        if q in ("Nmin", "Omin", "Nmax", "Omax") and not check_block:
            continue
        if q.startswith("min") and not min_volume:
            continue
        if q.startswith("Yh") and not yh:
            continue
        if q == "Cn" and not cn:
            continue
        # [X, Y, k] -> level-major, cell index Y * d1 + X
        out[q] = np.ascontiguousarray(np.transpose(tabs[q], (2, 1, 0))).reshape(-1)
    out["bottom"] = bottom.T.reshape(-1).copy()
    out["bottomXP"] = bXP.T.reshape(-1).copy()
    out["bottomYP"] = bYP.T.reshape(-1).copy()
    return out


def _min_volume(D, r, dx, Dx, Dim, tabs, k, h2):
    """script V16 :342-440 for level k, in its loop order."""
    d1, d2 = Dim
    dim = D.shape
    Np, Op = tabs["Np"], tabs["Op"]
    for ii in range(d1):
        for jj in range(d2):
            i1, i2 = ii * r, ii * r + r - 1
            j1, j2 = jj * r, jj * r + r - 1
            if i2 == dim[0] - 1:
                Lx = D[i1 + h2:i2 + 1, j1:j2 + 1]
            else:
                Lx = D[i1 + h2:i2 + h2 + 1, j1:j2 + 1]
            if j2 == dim[1] - 1:
                Ly = D[i1:i2 + 1, j1 + h2:j2 + 1]
            else:
                Ly = D[i1:i2 + 1, j1 + h2:j2 + h2 + 1]
            sN = [v for v in _rowsum(Lx) if v != 0]
            Azx = int((Lx > 0).sum())
            if sN:
                minVx = Dx * (min(sN) * dx)
                if minVx > 0:
                    if tabs["Nmax"][ii, jj, k] < Np[ii, jj, k]:
                        tabs["minVxp"][ii, jj, k] = 0.5 * minVx
                        if ii != d1 - 1:
                            tabs["minVxm"][ii + 1, jj, k] = 0.5 * minVx
                        else:
                            tabs["minVxm"][0, jj, k] = tabs["minVxp"][0, jj, k]
                    elif Azx / (r * r) > 0.2:
                        vp = tabs["Vxp"][ii, jj, k]
                        tabs["minVxp"][ii, jj, k] = max(min(0.5 * minVx, vp), 0.7 * vp)
                        if ii != d1 - 1:
                            vm = tabs["Vxm"][ii + 1, jj, k]
                            tabs["minVxm"][ii + 1, jj, k] = max(min(0.5 * minVx, vm), 0.7 * vm)
                        else:
                            tabs["minVxm"][0, jj, k] = tabs["minVxp"][0, jj, k]
                    else:
                        tabs["minVxp"][ii, jj, k] = tabs["Vxp"][ii, jj, k]
                        tabs["minVxm"][ii, jj, k] = tabs["Vxm"][ii, jj, k]
            else:
                tabs["minVxp"][ii, jj, k] = tabs["Vxp"][ii, jj, k]
                tabs["minVxm"][ii, jj, k] = tabs["Vxm"][ii, jj, k]
            sO = [v for v in _colsum(Ly) if v != 0]
            Azy = int((Ly > 0).sum())
            if sO:
                minVy = Dx * (min(sO) * dx)
                if minVy > 0:
                    if tabs["Omax"][ii, jj, k] < Op[ii, jj, k]:
                        tabs["minVyp"][ii, jj, k] = 0.5 * minVy
                        if jj != d2 - 1:
                            tabs["minVym"][ii, jj + 1, k] = 0.5 * minVy
                        else:
                            tabs["minVym"][ii, 0, k] = tabs["minVyp"][ii, 0, k]
                    elif Azy / (r * r) > 0.2:
                        vp = tabs["Vyp"][ii, jj, k]
                        tabs["minVyp"][ii, jj, k] = max(min(0.5 * minVy, vp), 0.7 * vp)
                        if jj != d2 - 1:
                            vm = tabs["Vym"][ii, jj + 1, k]
                            tabs["minVym"][ii, jj + 1, k] = max(min(0.5 * minVy, vm), 0.7 * vm)
                        else:
                            tabs["minVym"][ii, 0, k] = tabs["minVyp"][ii, 0, k]
                    else:
                        tabs["minVyp"][ii, jj, k] = tabs["Vyp"][ii, jj, k]
                        tabs["minVym"][ii, jj, k] = tabs["Vym"][ii, jj, k]
            else:
                tabs["minVyp"][ii, jj, k] = tabs["Vyp"][ii, jj, k]
                tabs["minVym"][ii, jj, k] = tabs["Vym"][ii, jj, k]
