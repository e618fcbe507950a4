"""Synthetic scenario builder for the free-surface step.

A scenario is what FrehdC reads from data.dat, bath.dat and the boundary
files (FREHD.c:30-63, bathymetry.c:20-57, initialize.c:345-378): grid and
physics settings, a bed elevation per cell, boundary series sampled once per
step, and the initial fields.  The builders here make the workloads the
benchmarks and tests use; every random draw is seeded.

Unspecified physics settings take the values of the reference's shipped
sample (FrehdC_v42/data.dat:67-87): eddy viscosity and diffusivity 1e-4,
thin-layer drag on (0.1 m, CD cap 1.0), minimum depth and waterfall
threshold 1e-5 m, advection CFL limiter 0.5/0.7.

Orientation: the reference applies tides only on the y boundaries
(YP = top row, configuration.c:143-148).  A "west" tidal boundary in a
workload description is mapped to the YP row, with the bed sloping along y.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import abi

TIDE_PERIOD = 44712.0
# the tidal (YP) boundary of the marsh workloads is an open-water inlet: the
# last rows are at least this deep, so the boundary cells never dry (a dry
# tidal cell with a near-zero face area makes the reference's boundary
# velocity flux/area blow up)
INLET_ROWS = 2
INLET_BED = -0.9


@dataclasses.dataclass
class Scenario:
    cfg: dict
    bottomZ: np.ndarray            # (N,) raw bed elevation, m about MSL
    tideP: np.ndarray              # (Nt+1,)
    tideM: np.ndarray
    inflow: np.ndarray
    tidalPS: np.ndarray
    tidalMS: np.ndarray
    tideLocP: np.ndarray           # int32 cell indices
    tideLocM: np.ndarray
    inflowLoc: np.ndarray
    manningCell: np.ndarray | None = None
    surf0: float = 0.0
    U0: float = 0.0
    V0: float = 0.0
    S0: float = 0.0
    S0cell: np.ndarray | None = None   # (nScalar*N,)
    name: str = ""
    surf0cell: np.ndarray | None = None  # (N,) per-cell initial fields (surfic/uuic/vvic.dat)
    U0cell: np.ndarray | None = None
    V0cell: np.ndarray | None = None
    subgrid: dict | None = None        # edgeX, edgeY and level tables (subgrid.py)
    windspd: np.ndarray | None = None  # (Nt+1,) with cfg useWind = 1
    winddir: np.ndarray | None = None

    @property
    def nx(self):
        return self.cfg["nx"]

    @property
    def ny(self):
        return self.cfg["ny"]

    @property
    def N(self):
        return self.nx * self.ny

    @property
    def Ng(self):
        return (self.nx + 2) * (self.ny + 2)

    # ---- ctypes views (the scenario object keeps the arrays alive) ----------
    def c_config(self, **override) -> abi.FrehdConfig:
        c = abi.FrehdConfig()
        vals = dict(self.cfg)
        vals.update(override)
        for k, v in vals.items():
            setattr(c, k, v)
        return c

    def c_forcing(self) -> abi.FrehdForcing:
        f = abi.FrehdForcing()
        f.tideLocP = abi.iptr(self.tideLocP)
        f.tideLocM = abi.iptr(self.tideLocM)
        f.inflowLoc = abi.iptr(self.inflowLoc)
        f.tideP = abi.dptr(self.tideP)
        f.tideM = abi.dptr(self.tideM)
        f.inflow = abi.dptr(self.inflow)
        f.tidalPS = abi.dptr(self.tidalPS)
        f.tidalMS = abi.dptr(self.tidalMS)
        f.manningCell = abi.dptr(self.manningCell)
        f.windspd = abi.dptr(self.windspd)
        f.winddir = abi.dptr(self.winddir)
        return f

    def c_ic(self) -> abi.FrehdIC:
        ic = abi.FrehdIC()
        ic.bottomZ = abi.dptr(self.bottomZ)
        ic.surf0, ic.U0, ic.V0, ic.S0 = self.surf0, self.U0, self.V0, self.S0
        ic.S0cell = abi.dptr(self.S0cell)
        ic.surf0cell = abi.dptr(self.surf0cell)
        ic.U0cell = abi.dptr(self.U0cell)
        ic.V0cell = abi.dptr(self.V0cell)
        if self.subgrid is not None:
            sg = abi.FrehdSubgrid()
            for k in ("edgeX", "edgeY", *abi.SUB_TABLES, *abi.DRAG_TABLES, *abi.MINA_TABLES):
                a = self.subgrid.get(k)
                setattr(sg, k, abi.dptr(a) if a is not None else None)
            if "fine" in self.subgrid:
                sg.fine = abi.dptr(self.subgrid["fine"])
                sg.fineR = int(self.subgrid["fineR"])
                sg.dxf = float(self.subgrid["dxf"])
            self._sg = sg
            ic.sub = C.pointer(sg)
        return ic


def default_config(nx, ny, dx, dt, Nt, **kw) -> dict:
    cfg = dict(
        nx=nx, ny=ny, Nt=Nt, maxIter=10000, dx=float(dx), dy=float(dx), dt=float(dt), g=9.81,
        NUx=1e-4, NUy=1e-4, CDx=0.01, CDy=0.01, manningN=0.025, CDnotN=0, useThinLayer=1,
        hD=0.1, CDmax=1.0, minDepth=1e-5, wtfh=1e-5, CFLl=0.5, CFLh=0.7, eps=1e-10,
        precond=abi.PRECOND_JACOBI, precision=abi.PREC_FP64, bcType=1, tideLocLengthP=nx,
        tideLocLengthM=0, inflowLocLength=0, useScalar=1, scalarAdv=1, useConstTidePS=1,
        useConstTideMS=1, nScalar=1, tidePS=0.0, tideMS=0.0, inflowS=0.0, Kx=1e-4, Ky=1e-4,
        offset=0.0, epsInner=1e-5, maxRefine=8, warmStart=0, useSubgrid=0, Ne=0, surfmin=0.0, dsurf=0.01,
        subR=1.0, subgridOnTheFly=0, useWind=0, northAngle=0.0, rhoa=1.225, Cw=1.3e-3, CwT=1.0,
        useSubDrag=0, useminA=0, useminV=1, beta=0.0,
    )
    cfg.update(kw)
    return cfg


def offset_for(bed: np.ndarray) -> float:
    """Bath.offset: -min(bed) when the bed dips below zero (bathymetry.c:79-83)."""
    zmin = float(np.min(bed))
    return 0.0 if zmin >= 0 else -zmin


def tide_series(Nt, dt, amp, period=TIDE_PERIOD, phase=0.0):
    t = np.arange(Nt + 1, dtype=np.float64) * dt
    return amp * np.sin(2.0 * math.pi * t / period + phase)


def fbm_surface(nx, ny, hurst, rng, lo, hi, uniform=False):
    """Fractional-Brownian surface by spectral synthesis (power ~ k^-(2H+2)),
    rescaled to [lo, hi].  uniform=True maps it monotonically (by rank) so the
    elevations spread uniformly over [lo, hi], keeping the spatial pattern."""
    kx = np.fft.fftfreq(nx)[None, :]
    ky = np.fft.fftfreq(ny)[:, None]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    phase = rng.uniform(0, 2 * np.pi, size=(ny, nx))
    noise = rng.normal(size=(ny, nx))
    spec = amp * noise * np.exp(1j * phase)
    f = np.real(np.fft.ifft2(spec))
    if uniform:
        r = np.empty(f.size)
        r[np.argsort(f, axis=None, kind="stable")] = (np.arange(f.size) + 0.5) / f.size
        f = r.reshape(f.shape)
    else:
        f = (f - f.min()) / max(f.max() - f.min(), 1e-300)
    return lo + (hi - lo) * f


def _finish(cfg, bed, tideP, scalar0=None, manning=None, inflowLoc=None, inflow=None,
            surf0=0.0, name=""):
    nx, ny, Nt = cfg["nx"], cfg["ny"], cfg["Nt"]
    N = nx * ny
    bed = np.ascontiguousarray(bed, dtype=np.float64).reshape(N)
    cfg["offset"] = offset_for(bed)
    z = np.zeros(Nt + 1)
    tideLocP = np.arange(nx * (ny - 1), nx * ny, dtype=np.int32)[: cfg["tideLocLengthP"]]
    if inflowLoc is None:
        inflowLoc = np.zeros(0, dtype=np.int32)
    cfg["inflowLocLength"] = int(len(inflowLoc))
    return Scenario(
        cfg=cfg, bottomZ=bed, tideP=np.ascontiguousarray(tideP), tideM=z.copy(),
        inflow=np.ascontiguousarray(inflow if inflow is not None else z.copy()),
        tidalPS=z.copy(), tidalMS=z.copy(), tideLocP=np.ascontiguousarray(tideLocP),
        tideLocM=np.zeros(0, dtype=np.int32), inflowLoc=np.ascontiguousarray(inflowLoc, dtype=np.int32),
        manningCell=None if manning is None else np.ascontiguousarray(manning, dtype=np.float64),
        surf0=surf0, S0cell=None if scalar0 is None else np.ascontiguousarray(scalar0, dtype=np.float64),
        name=name)


def basin(nx=128, ny=128, dx=10.0, dt=1.0, Nt=3600, seed=1, amp=0.2, bumps=20):
    """Rectangular basin, bed sloping from 3.0 m depth at the tidal boundary to
    0.5 m at the far wall plus seeded Gaussian bumps (0.3 m, sigma 30 m);
    Manning 0.025; Gaussian scalar plume (peak 1, sigma 50 m) at the centre;
    Jacobi PCG to 1e-10."""
    rng = np.random.default_rng(seed)
    y = np.arange(ny)[:, None] * dx
    x = np.arange(nx)[None, :] * dx
    s = (ny - 1 - np.arange(ny)[:, None]) / max(ny - 1, 1)     # 0 at the tidal row
    bed = (-3.0 + 2.5 * s) * np.ones((1, nx))
    cx = rng.uniform(0, nx * dx, size=bumps)
    cy = rng.uniform(0, ny * dx, size=bumps)
    for a, b in zip(cx, cy):
        bed = bed + 0.3 * np.exp(-((x - a) ** 2 + (y - b) ** 2) / (2 * 30.0 ** 2))
    cfg = default_config(nx, ny, dx, dt, Nt, manningN=0.025, eps=1e-10)
    xc, yc = nx * dx / 2, ny * dx / 2
    plume = np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / (2 * 50.0 ** 2)) * np.ones((ny, nx))
    return _finish(cfg, bed, tide_series(Nt, dt, amp), scalar0=plume.reshape(-1),
                   name=f"basin{nx}x{ny}")


def marsh(n=1024, dx=2.0, dt=0.5, Nt=2000, seed=2, hurst=0.7, amp=0.6):
    """Marsh with wetting and drying: fBm bed (Hurst 0.7) spread uniformly over
    [-1.0, +1.5] m, tide 0.6 m, per-cell Manning n ~ U(0.02, 0.04), scalar 0
    inside with 1 at the tidal boundary."""
    rng = np.random.default_rng(seed)
    bed = fbm_surface(n, n, hurst, rng, -1.0, 1.5, uniform=True)
    bed[-INLET_ROWS:, :] = np.minimum(bed[-INLET_ROWS:, :], INLET_BED)
    manning = np.random.default_rng(seed + 1).uniform(0.02, 0.04, size=n * n)
    cfg = default_config(n, n, dx, dt, Nt, useConstTidePS=1, tidePS=1.0, eps=1e-10)
    return _finish(cfg, bed, tide_series(Nt, dt, amp), scalar0=np.zeros(n * n),
                   manning=manning, name=f"marsh{n}")


def channel(n=4096, dx=1.0, dt=0.5, Nt=500, seed=4, hurst=0.8, amp=0.5, q_river=50.0,
            nscalar=3):
    """Large domain: fBm bed (Hurst 0.8) with a carved meandering channel
    (2 m deep, 20 cells wide), tide 0.5 m on the YP row, river inflow of
    q_river m^3/s spread over the channel cells of row 0, three scalars."""
    rng = np.random.default_rng(seed)
    bed = fbm_surface(n, n, hurst, rng, -0.5, 1.5)
    jj = np.arange(n)
    centre = n / 2 + 0.2 * n * np.sin(2 * np.pi * jj / n * 2.0)
    ii = np.arange(n)[None, :]
    inside = np.abs(ii - centre[:, None]) < 10
    bed = np.where(inside, bed - 2.0, bed)
    cfg = default_config(n, n, dx, dt, Nt, eps=1e-10, nScalar=nscalar)
    row0 = np.nonzero(inside[0])[0].astype(np.int32)
    inflowLoc = row0.copy()                     # cells of row j = 0
    inflow = np.full(Nt + 1, q_river)
    N = n * n
    s0 = np.concatenate([np.full(N, 35.0 * (q / max(nscalar - 1, 1))) for q in range(nscalar)])
    return _finish(cfg, bed, tide_series(Nt, dt, amp), scalar0=s0, inflowLoc=inflowLoc,
                   inflow=inflow, name=f"channel{n}")


def fbm_pixels(nx, ny, hurst, rng, lo, hi, workers=None):
    """fBm surface as fbm_surface (rescaled to [lo, hi]) for pixel grids of
    10^8 points and more: single-precision spectrum, multithreaded inverse
    FFT (scipy.fft), float64 result."""
    import os
    import scipy.fft
    kx = np.fft.fftfreq(nx).astype(np.float32)[None, :]
    ky = np.fft.fftfreq(ny).astype(np.float32)[:, None]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** np.float32(-(hurst + 1.0))
    amp[0, 0] = 0.0
    del k
    spec = np.empty((ny, nx), dtype=np.complex64)
    for j0 in range(0, ny, 1024):                     # row blocks: bounded temporaries
        j1 = min(ny, j0 + 1024)
        phase = rng.uniform(0, 2 * np.pi, size=(j1 - j0, nx)).astype(np.float32)
        noise = rng.normal(size=(j1 - j0, nx)).astype(np.float32)
        spec[j0:j1] = amp[j0:j1] * noise * (np.cos(phase) + 1j * np.sin(phase))
    del amp
    f = scipy.fft.ifft2(spec, workers=workers or min(16, os.cpu_count() or 1), overwrite_x=True).real
    del spec
    fmin, fmax = float(f.min()), float(f.max())
    out = np.empty((ny, nx), dtype=np.float64)
    np.subtract(f, fmin, out=out)
    out *= (hi - lo) / max(fmax - fmin, 1e-30)
    out += lo
    return out


def channel_subgrid(n=4096, r=4, dx=1.0, dt=0.5, Nt=500, seed=4, hurst=0.8, amp=0.5, q_river=50.0,
                    nscalar=3, lo=-1.5, hi=1.0, width=20, depth=2.0, dsurf=0.01, on_the_fly=True):
    """The large domain on the subgrid model: r x r pixels per cell of an fBm
    bed (Hurst 0.8) in [lo, hi] with a meandering channel (depth m deep,
    `width` cells wide) carved at pixel resolution; tide amp on the YP row,
    river inflow q_river m^3/s spread over the channel cells of row 0; three
    scalars: a salinity-like 0..35 gradient along y, a temperature-like
    10..30 gradient along x, a tracer 0/1 patch; boundary scalar 1 at the
    tide.  Levels span [lo - depth, hi] every dsurf; the device keeps the
    pixels and evaluates levels per lookup (on_the_fly) since the tables of
    a 4096^2 grid would need 10 x Ne x N values."""
    rng = np.random.default_rng(seed)
    fy = fx = n * r
    fine = fbm_pixels(fx, fy, hurst, rng, lo, hi)
    yp = np.arange(fy, dtype=np.float64)
    centre = fx / 2 + 0.2 * fx * np.sin(2 * np.pi * yp / fy * 2.0)
    xp = np.arange(fx, dtype=np.float64)
    for y0 in range(0, fy, 2048):
        y1 = min(fy, y0 + 2048)
        inside = np.abs((xp[None, :] + 0.5) - centre[y0:y1, None]) < width * r / 2
        fine[y0:y1][inside] -= depth
    cfg = default_config(n, n, dx, dt, Nt, eps=1e-10, nScalar=nscalar, useConstTidePS=1, tidePS=1.0)
    N = n * n
    jj, ii = np.divmod(np.arange(N), n)
    s0 = [35.0 * jj / (n - 1), 10.0 + 20.0 * ii / (n - 1),
          ((np.abs(ii - n // 2) < n // 8) & (np.abs(jj - n // 2) < n // 8)).astype(np.float64)]
    s0 = np.concatenate([s0[q % 3] for q in range(nscalar)])
    # inflow cells: row 0 cells whose middle pixel lies in the channel
    mid = np.abs((xp[r // 2::r] + 0.5) - centre[r // 2]) < width * r / 2
    row0 = np.nonzero(mid)[0].astype(np.int32)
    inflow = np.full(Nt + 1, q_river / len(row0))
    sc = _finish(cfg, np.zeros(N), tide_series(Nt, dt, amp), scalar0=s0, inflowLoc=row0, inflow=inflow,
                 name=f"channel{n}")
    sc = with_subgrid(sc, fine, r, lo - depth, hi, dsurf, on_the_fly=on_the_fly)
    sc.name = f"channel{n}_sub{r}"
    return sc


def ensemble_member(m, n=2048, dx=1.0, dt=0.5, Nt=1000, seed=5, hurst=0.7, base=None, period=TIDE_PERIOD):
    """Ensemble member m: shared fBm base plus a seeded N(0, 0.05) perturbation,
    Manning n ~ U(0.015, 0.04) and tide amplitude ~ U(0.3, 0.8) per member
    (tidal period `period`)."""
    if base is None:
        base = fbm_surface(n, n, hurst, np.random.default_rng(seed), -1.0, 1.5)
    r = np.random.default_rng(100 + m)
    bed = base + r.normal(0.0, 0.05, size=base.shape)
    bed[-INLET_ROWS:, :] = np.minimum(bed[-INLET_ROWS:, :], INLET_BED)   # always-wet tidal inlet
    cfg = default_config(n, n, dx, dt, Nt, manningN=float(r.uniform(0.015, 0.04)), eps=1e-10)
    return _finish(cfg, bed, tide_series(Nt, dt, float(r.uniform(0.3, 0.8)), period=period),
                   scalar0=np.zeros(n * n), name=f"member{m}")


def serving_requests(count=512, rate=20.0, seed=6, lo=64, hi=1024, dx=2.0, dt=0.5,
                     steps_lo=200, steps_hi=2000, subgrid_r=0, on_the_fly=False):
    """Poisson arrivals of scenarios with log-uniform sizes in [lo^2, hi^2]
    (rounded to multiples of 64 per side), fBm bed with Hurst ~ U(0.6, 0.9) in
    [-1.5, +1] m (an always-wet inlet at the tidal rows), tide ~ U(0.1, 0.7) m,
    Manning ~ U(0.015, 0.04), 1-3 scalars, U(200, 2000) steps; with
    subgrid_r > 0 the bed is drawn at r x r pixels per cell and the tables
    are generated on the device (or, on_the_fly, the device keeps the pixels
    and evaluates the levels at each lookup: r^2 doubles per cell instead
    of 10 per level).  Returns [(arrival_s, steps, side, builder)]
    where builder() makes the Scenario on demand."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    out = []
    for q in range(count):
        side = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        side = int(max(lo, min(hi, round(side / 64) * 64)))
        hurst = float(rng.uniform(0.6, 0.9))
        amp = float(rng.uniform(0.1, 0.7))
        man = float(rng.uniform(0.015, 0.04))
        nsc = int(rng.integers(1, 4))
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        sseed = int(rng.integers(1 << 30))

        def build(side=side, hurst=hurst, amp=amp, man=man, nsc=nsc, steps=steps, sseed=sseed):
            r = max(1, subgrid_r)
            bed = fbm_surface(side * r, side * r, hurst, np.random.default_rng(sseed), -1.5, 1.0)
            bed[-INLET_ROWS * r:, :] = np.minimum(bed[-INLET_ROWS * r:, :], INLET_BED)
            cfg = default_config(side, side, dx, dt, steps, manningN=man, nScalar=nsc, eps=1e-10)
            coarse = bed.reshape(side, r, side, r).min(axis=(1, 3)) if r > 1 else bed
            sc = _finish(cfg, coarse, tide_series(steps, dt, amp),
                         scalar0=np.zeros(nsc * side * side), name=f"req{side}")
            if subgrid_r > 0:
                sc = with_subgrid(sc, bed, r, -1.5, 1.0, 0.01, device_tables=True, on_the_fly=on_the_fly)
            return sc
        out.append((float(t[q]), steps, side, build))
    return out


def with_wind(scen: Scenario, speed: float = 8.0, gust: float = 3.0, direction: float = 30.0,
              veer: float = 40.0, northAngle: float = 10.0, period: float = 600.0) -> Scenario:
    """Switch on the wind stress (windTerm, nsfunctions.c:609-648): speed
    speed + gust sin(2 pi t / period) m/s from direction + veer sin(pi t /
    period) degrees east of north, per step (bc->windspd / winddir, Nt+1
    samples); Config.rhoa, Cw, CwT keep the scenario's values."""
    Nt, dt = scen.cfg["Nt"], scen.cfg["dt"]
    t = np.arange(Nt + 1, dtype=np.float64) * dt
    scen.windspd = np.ascontiguousarray(speed + gust * np.sin(2.0 * math.pi * t / period))
    scen.winddir = np.ascontiguousarray(direction + veer * np.sin(math.pi * t / period))
    scen.cfg.update(useWind=1, northAngle=float(northAngle))
    scen.name = scen.name + "_wind"
    return scen


def with_subdrag(scen: Scenario, model: int, seed: int = 11) -> Scenario:
    """Switch a table-based subgrid scenario to a subgrid drag model
    (updateCDSub, subgrid.c:1481-1519): 2 (Volp 2013) with YhX/YhY tables, 3
    or 4 (Casas 2010) with a Cn table.  The tables are synthetic but shaped
    like the real ones, level-major [k*N + c]: Yh falls with the water depth
    above the level (a depth factor that the CD cap meets on shallow levels),
    Cn is a per-cell roughness in [0.015, 0.05] that grows on dry levels."""
    sg, k = scen.subgrid, scen.cfg
    assert sg is not None and "V" in sg, "needs level tables (not device tables)"
    N, Ne = scen.N, k["Ne"]
    rng = np.random.default_rng(seed)
    lev = (k["surfmin"] + k["dsurf"] * np.arange(Ne))[:, None]
    depth = np.maximum(lev - scen.bottomZ[None, :], 0.0)
    if model == 2:
        for name in ("YhX", "YhY"):
            f = rng.uniform(0.8, 1.25, size=N)[None, :]
            yh = np.where(depth > 0, f / np.cbrt(np.maximum(depth, 0.05)), 0.0)
            sg[name] = np.ascontiguousarray(yh.reshape(-1))
    elif model in (3, 4):
        base = rng.uniform(0.015, 0.04, size=N)[None, :]
        sg["Cn"] = np.ascontiguousarray(np.where(depth > 0, base, base + 0.01).reshape(-1))
    else:
        raise ValueError("useSubDrag 2, 3 or 4")
    k["useSubDrag"] = int(model)
    scen.name += f"_subdrag{model}"
    return scen


def with_minarea(scen: Scenario, beta: float = 0.5, useminV: int = 1, seed: int = 13) -> Scenario:
    """Switch on the minimum face areas (useminA = 1, Config.beta) and the
    effective volumes (useminV) of combineSubArea (subgrid.c:360-414) on a
    table-based subgrid scenario.  The tables are synthetic, shaped like the
    real ones (level-major): Nmax/Omax a per-cell fraction in [0.6, 1.1] of the
    face-area tables Np/Op, Nmin/Omin a fraction in [0.2, 0.7] of Nmax/Omax,
    the minimum half-cell volumes a fraction in [0.5, 0.9] of Vxp.. Vym."""
    sg, k = scen.subgrid, scen.cfg
    assert sg is not None and "V" in sg, "needs level tables (not device tables)"
    N, Ne = scen.N, k["Ne"]
    rng = np.random.default_rng(seed)

    def frac(lo, hi):
        return np.tile(rng.uniform(lo, hi, size=N), Ne)

    for mx, mn, src in (("Nmax", "Nmin", "Np"), ("Omax", "Omin", "Op")):
        sg[mx] = np.ascontiguousarray(sg[src] * frac(0.6, 1.1))
        sg[mn] = np.ascontiguousarray(sg[mx] * frac(0.2, 0.7))
    for m, src in (("minVxp", "Vxp"), ("minVxm", "Vxm"), ("minVyp", "Vyp"), ("minVym", "Vym")):
        sg[m] = np.ascontiguousarray(sg[src] * frac(0.5, 0.9))
    k.update(useminA=1, useminV=int(useminV), beta=float(beta))
    scen.name += f"_minA{useminV}"
    return scen


def with_generated_tables(scen: Scenario, fine: np.ndarray, useminA: int = 1, useminV: int = 1,
                          beta: float = 0.5, subdrag: int = 0, check_block: bool = True) -> Scenario:
    """Replace a table-based subgrid scenario's tables by the generator's
    with the reference's block check (frehdc_amd/subgrid.py generate():
    blocked faces, minimum / maximum face areas, minimum half-cell volumes,
    and the Volp (subdrag 2) or Casas (3, 4) drag tables), from the pixels
    `fine` the scenario was built on, and switch on the options that read
    them (useminA, useminV, Config.beta, useSubDrag)."""
    from . import subgrid as sgm
    sg, k = scen.subgrid, scen.cfg
    assert sg is not None and "V" in sg, "needs level tables (not device tables)"
    r = int(round(k["subR"]))
    t = sgm.generate(fine, r, k["dx"] / r, k["surfmin"], k["dsurf"], k["Ne"], check_block=check_block,
                     min_volume=True, yh=subdrag == 2, cn=subdrag in (3, 4))
    for name, v in t.items():
        if name not in ("bottom", "bottomXP", "bottomYP"):
            sg[name] = np.ascontiguousarray(v)
    k.update(useminA=int(useminA), useminV=int(useminV), beta=float(beta))
    if subdrag:
        k["useSubDrag"] = int(subdrag)
    scen.name += f"_gen{'B' if check_block else ''}_minA{useminA}V{useminV}" + (f"_drag{subdrag}" if subdrag else "")
    return scen


def marsh_subgrid_generated(n=28, r=4, Nt=300, period=1200.0, useminA=1, useminV=1, beta=0.5, subdrag=0,
                            manning_cell=True, seed=2):
    """The subgrid marsh on the generator's block-checked tables (see
    with_generated_tables)."""
    sc = marsh_subgrid(n=n, r=r, Nt=Nt, period=period, manning_cell=manning_cell, seed=seed)
    return with_generated_tables(sc, fine_marsh(n, r, seed=seed), useminA=useminA, useminV=useminV, beta=beta,
                                 subdrag=subdrag)


def with_subgrid(scen: Scenario, fine: np.ndarray, r: int, surfmin: float, surfmax: float,
                 dsurf: float = 0.01, device_tables: bool = False, on_the_fly: bool = False) -> Scenario:
    """Turn a scenario into its subgrid form: `fine` holds the pixel bed
    elevations [ny*r, nx*r]; the coarse bed is each block's minimum
    (bottom.dat), face bottoms and level tables come from subgrid.tables(),
    or, with device_tables, the fine bathymetry is handed to the device
    library, which generates the tables itself (with on_the_fly it keeps the
    pixels and evaluates the levels at each lookup instead)."""
    device_tables = device_tables or on_the_fly
    from . import subgrid as sgm
    nx, ny = scen.nx, scen.ny
    assert fine.shape == (ny * r, nx * r)
    bottom, edgeX, edgeY = sgm.coarse_geometry(fine, r)
    Ne = int(round((surfmax - surfmin) / dsurf))
    dxf = scen.cfg["dx"] / r
    if device_tables:
        tabs = {"fine": np.ascontiguousarray(fine, dtype=np.float64), "fineR": r, "dxf": dxf}
    else:
        tabs = sgm.tables(fine, r, dxf, surfmin, dsurf, Ne)
    scen.bottomZ = np.ascontiguousarray(bottom.reshape(-1))
    scen.cfg["offset"] = offset_for(scen.bottomZ)
    scen.cfg.update(useSubgrid=1, Ne=Ne, surfmin=float(surfmin), dsurf=float(dsurf), subR=float(r),
                    subgridOnTheFly=1 if on_the_fly else 0)
    tabs["edgeX"] = np.ascontiguousarray(edgeX.reshape(-1))
    tabs["edgeY"] = np.ascontiguousarray(edgeY.reshape(-1))
    scen.subgrid = tabs
    scen.name += f"_sub{r}"
    return scen


def fine_basin(nx, ny, r, dx=10.0, seed=1, bumps=20):
    """Pixel bed of the basin workload: slope from -3.0 m at the tidal row to
    -0.5 m, plus Gaussian bumps (0.3 m, sigma 30 m) and pixel-scale seeded
    roughness (0.05 m) so the subgrid areas differ from the plain grid."""
    rng = np.random.default_rng(seed)
    dxf = dx / r
    fy, fx = ny * r, nx * r
    y = (np.arange(fy)[:, None] + 0.5) * dxf
    x = (np.arange(fx)[None, :] + 0.5) * dxf
    s = (ny * dx - y) / (ny * dx)
    bed = (-3.0 + 2.5 * s) * np.ones((1, fx))
    cx = rng.uniform(0, nx * dx, size=bumps)
    cy = rng.uniform(0, ny * dx, size=bumps)
    for a, b in zip(cx, cy):
        bed = bed + 0.3 * np.exp(-((x - a) ** 2 + (y - b) ** 2) / (2 * 30.0 ** 2))
    return bed + rng.normal(0.0, 0.05, size=bed.shape)


def basin_subgrid(nx=128, ny=128, r=4, dx=10.0, dt=1.0, Nt=3600, seed=1, amp=0.2, dsurf=0.01,
                  device_tables=False, on_the_fly=False):
    """The basin workload on the subgrid model: r x r pixels per cell of
    fine_basin's bed (slope, bumps, pixel roughness), levels from 0.5 m
    below the deepest pixel to 0.5 m above the highest tide every dsurf;
    staggeredV = 1, useminA = 0, useminV = 1, useSubDrag = 0 (the
    scenario defaults)."""
    sc = basin(nx=nx, ny=ny, dx=dx, dt=dt, Nt=Nt, seed=seed, amp=amp)
    fine = fine_basin(nx, ny, r, dx=dx, seed=seed)
    lo = round(math.floor((float(fine.min()) - 0.5) / dsurf) * dsurf, 9)
    hi = round(math.ceil((amp + 0.5) / dsurf) * dsurf, 9)
    sc = with_subgrid(sc, fine, r, lo, hi, dsurf, device_tables=device_tables, on_the_fly=on_the_fly)
    sc.name = f"basin{nx}x{ny}_sub{r}"
    return sc


def fine_marsh(n, r, seed=2, hurst=0.7, lo=-1.0, hi=1.5):
    """Pixel bed of the marsh: fBm (Hurst 0.7) at pixel resolution spread
    uniformly over [lo, hi]."""
    rng = np.random.default_rng(seed)
    f = fbm_surface(n * r, n * r, hurst, rng, lo, hi, uniform=True)
    f[-INLET_ROWS * r:, :] = np.minimum(f[-INLET_ROWS * r:, :], INLET_BED)
    return f


def marsh_subgrid(n=48, r=8, dx=2.0, dt=0.5, Nt=400, amp=0.6, period=TIDE_PERIOD, seed=2,
                  dsurf=0.01, manning_cell=True, device_tables=False, on_the_fly=False):
    """The marsh workload on the subgrid model: r x r pixels per cell, tables
    from surfmin = -1.0 m to surfmax = +1.5 m every dsurf (generated on the
    device with device_tables, evaluated per lookup with on_the_fly)."""
    sc = marsh(n=n, dx=dx, dt=dt, Nt=Nt, seed=seed, amp=amp)
    if period != TIDE_PERIOD:
        sc.tideP[:] = tide_series(Nt, dt, amp, period=period)
    if not manning_cell:
        sc.manningCell = None
    fine = fine_marsh(n, r, seed=seed)
    sc = with_subgrid(sc, fine, r, -1.0, 1.5, dsurf, device_tables=device_tables, on_the_fly=on_the_fly)
    sc.name = f"marsh{n}_sub{r}"
    return sc
