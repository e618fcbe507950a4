"""The device library builds for the CPU-only container and exports exactly
the C ABI declared in include/frehd_gpu.h, frehd_io.h and frehd_subgrid.h;
the ctypes mirror matches the
compiled struct sizes.  No device call is made here."""
import ctypes as C
import os
import re
import subprocess

import pytest

from frehdc_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the headers whose functions libfrehd_gpu.so implements (frehd_seam.h's are
# the seam library's)
HEADERS = [os.path.join(ROOT, "include", h) for h in ("frehd_gpu.h", "frehd_io.h", "frehd_subgrid.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(frehd_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(abi.LIB_PATH):
        from frehdc_amd import build
        build.build()
    return abi.load_library()


def test_exports_every_declared_function(lib):
    names = declared_functions()
    assert len(names) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(frehd_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)


def test_struct_sizes_match(lib):
    sizes = (C.c_int64 * 6)()
    assert lib.frehd_abi_sizes(sizes) == 0
    ours = [C.sizeof(abi.FrehdConfig), C.sizeof(abi.FrehdForcing), C.sizeof(abi.FrehdIC),
            C.sizeof(abi.FrehdFields), C.sizeof(abi.FrehdStats), C.sizeof(abi.FrehdBudget)]
    assert list(sizes) == ours


def test_error_paths_need_no_device(lib):
    # argument checks run before any device call
    assert lib.frehd_dev_step(None, 1) == abi.FREHD_EINVAL
    assert b"NULL" in lib.frehd_last_error()
    cfg = abi.FrehdConfig(nx=2, ny=2, Nt=1, dx=1.0, dy=1.0, dt=1.0)
    b = C.c_int64()
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == abi.FREHD_EINVAL
    # SSOR preconditions the FP64 solve (one more vector: z, but not the five
    # per-tile partials of the one-kernel Jacobi solve, 2 x 5 x (64 + 8)
    # doubles here); not the mixed path
    cfg = abi.FrehdConfig(nx=64, ny=64, Nt=10, dx=1.0, dy=1.0, dt=1.0, bcType=1, precond=1)
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == 0
    with_z = b.value
    cfg = abi.FrehdConfig(nx=64, ny=64, Nt=10, dx=1.0, dy=1.0, dt=1.0, bcType=1, precond=0)
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == 0
    assert with_z - b.value >= 64 * 64 * 8 - 2 * 5 * 72 * 8 - 2 * 256
    cfg = abi.FrehdConfig(nx=64, ny=64, Nt=10, dx=1.0, dy=1.0, dt=1.0, bcType=1, precond=1, precision=1,
                          epsInner=1e-5, maxRefine=4)
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == abi.FREHD_EINVAL
    assert b"SSOR" in lib.frehd_last_error()
    cfg = abi.FrehdConfig(nx=64, ny=64, Nt=10, dx=1.0, dy=1.0, dt=1.0, bcType=1, precond=2)
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == abi.FREHD_EINVAL
    # subgrid drag: model 1 reads curvature tables the reference never fills;
    # the models need level tables, not the on-the-fly evaluation
    sub = dict(nx=64, ny=64, Nt=10, dx=1.0, dy=1.0, dt=1.0, bcType=1, useSubgrid=1, Ne=10, dsurf=0.01,
               subR=2.0)
    cfg = abi.FrehdConfig(useSubDrag=1, **sub)
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == abi.FREHD_EINVAL
    assert b"useSubDrag" in lib.frehd_last_error()
    cfg = abi.FrehdConfig(useSubDrag=2, subgridOnTheFly=1, **sub)
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == abi.FREHD_EINVAL
    cfg = abi.FrehdConfig(useSubDrag=2, **sub)
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == 0
    base = b.value
    cfg = abi.FrehdConfig(useSubDrag=0, **sub)
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == 0
    assert base - b.value >= 2 * 10 * 64 * 64 * 8     # the two Yh tables


def test_footprint_scales(lib):
    from frehdc_amd import cases
    sc = cases.basin(nx=256, ny=256, Nt=10)
    cfg = sc.c_config()
    b = C.c_int64()
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == 0
    per_cell = b.value / sc.N
    assert 200 < per_cell < 600, per_cell


def test_on_the_fly_footprint():
    """Without tables the handle holds r^2 pixels per cell instead of
    10 x Ne table values."""
    from frehdc_amd import cases
    sc = cases.marsh_subgrid(n=64, r=4, Nt=10, device_tables=True)
    fly = dict(sc.cfg, subgridOnTheFly=1)
    lib = abi.load_library()
    import ctypes as C
    t, f = C.c_int64(), C.c_int64()
    abi.check(lib.frehd_dev_footprint(C.byref(sc.c_config()), C.byref(t)))
    cf = abi.FrehdConfig()
    for k, v in fly.items():
        setattr(cf, k, v)
    abi.check(lib.frehd_dev_footprint(C.byref(cf), C.byref(f)))
    N = sc.N
    assert t.value - f.value == pytest.approx(8 * N * (10 * sc.cfg["Ne"] - 16), abs=4096)


def test_mixed_precision_size_limit(lib):
    # the inner FP32 solve addresses its vectors with 32-bit byte offsets
    b = C.c_int64()
    mixed = dict(Nt=10, dx=1.0, dy=1.0, dt=1.0, bcType=1, precision=abi.PREC_MIXED, epsInner=1e-6, maxRefine=8)
    cfg = abi.FrehdConfig(nx=16384, ny=32768, **mixed)          # exactly 2^29 cells
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == abi.FREHD_EINVAL
    assert b"2^29" in lib.frehd_last_error()
    cfg = abi.FrehdConfig(nx=16384, ny=32767, **mixed)
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == 0
    # the one-kernel FP64 Jacobi solve does the same with its FP64 vectors
    # (2^28 cells at most); the SSOR path addresses them directly
    fp64 = dict(mixed, precision=abi.PREC_FP64)
    cfg = abi.FrehdConfig(nx=16384, ny=16385, **fp64)
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == abi.FREHD_EINVAL
    assert b"2^28" in lib.frehd_last_error()
    cfg = abi.FrehdConfig(nx=16384, ny=16384, **fp64)          # exactly 2^28 cells
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == 0
    cfg = abi.FrehdConfig(nx=16384, ny=32768, **dict(fp64, precond=abi.PRECOND_SSOR))
    assert lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)) == 0
