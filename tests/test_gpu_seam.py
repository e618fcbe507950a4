"""The reference-signature entry points (frehdc_amd/seam/libfrehd_seam.so,
compiled against FrehdC's headers) driven by the reference's own caller
code, on the reference's own Init state (oracle/_ref):

* oneCompleteStep (nssolve.h:10-11) inside SolveAll's loop body
  (nssolve.c:61-104: the LASpack objects built and destroyed around each
  call, the reference's monitorCFL after it) over 60 wet/dry steps must give,
  bit for bit, what the device gives resident (frehd_dev_step from the same
  state): each call hands the host state over and takes it back;
* SolveAllOnDevice, called as SolveAll is (FREHD.c:105), must give the
  resident run too, after DataOutput's in-place offset round trip of the
  surface at step 0 (fileio.c:65-66, 120-121) and with the last forcing
  sample repeated (the reference's series hold Nt samples, step Nt reads one
  more).
Both solve as the reference does, with SSOR-preconditioned CG."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from frehdc_amd import abi, build, cases
from frehdc_amd.device import DeviceScenario
import pyoracle as po

pytestmark = pytest.mark.gpu

STATE = ["surf", "surfOld", "uuXP", "vvYP", "uuYP", "vvXP", "depth", "depthXP", "depthYP", "EnXP", "EnYP",
         "DragXP", "DragYP", "CDXP", "CDYP", "Fuu", "Fvv", "cellV", "S", "Sm", "CFLx", "CFLy", "wtfXP", "wtfYP",
         "bottomZ", "bottomXP", "bottomYP"]
SUB = ["V", "Z", "Vx", "Vy", "Zx", "Zy", "Nx", "Oy"]

_seam = None


def seam():
    """The seam library, its reference-side symbols (DataOutput, SubOutput)
    taken from the reference build loaded first with global symbols."""
    global _seam
    if _seam is None:
        if not po.have_reference():
            pytest.skip("reference build absent")
        C.CDLL(po.REF_LIB, mode=C.RTLD_GLOBAL)
        abi.load_library()
        _seam = C.CDLL(build.SEAM_LIB)
        _seam.frehd_seam_release.argtypes = []
        _seam.frehd_seam_release.restype = None
    return _seam


def ref_state(r, sub):
    out = {n: r.raw(n) for n in STATE}
    if sub:
        out.update({n: r.raw(n) for n in SUB})
        out["ind"] = r.ind()
    return out


def assert_same(r, dev, sub, tag):
    names = STATE + (SUB if sub else [])
    got = dev.download(names + (["ind"] if sub else []))
    for n in names:
        a = r.raw(n)
        b = np.asarray(got[n]).reshape(-1)
        m = min(len(a), len(b))
        assert np.array_equal(a[:m].view(np.int64), b[:m].view(np.int64)), (tag, n)
    if sub:
        assert np.array_equal(r.ind(), got["ind"]), (tag, "ind")


CASES = {"marsh40": lambda: cases.marsh(n=40, dx=2.0, dt=0.5, Nt=120, amp=0.6),
         "marsh24_sub4": lambda: cases.marsh_subgrid(n=24, r=4, Nt=120, period=1200.0, manning_cell=False)}


def short_tide(sc):
    sc.manningCell = None                  # the reference's single Manning n
    sc.tideP[:] = cases.tide_series(sc.cfg["Nt"], sc.cfg["dt"], 0.6, period=1200.0)
    return sc


@pytest.mark.parametrize("name", list(CASES))
def test_one_complete_step_seam(name):
    lib = seam()
    sc = short_tide(CASES[name]())
    sub = sc.subgrid is not None
    print(f"case {name}: oneCompleteStep seam in SolveAll's loop vs resident, 60 steps", flush=True)
    r = po.Reference(sc)
    init = ref_state(r, sub)
    lib.frehd_seam_release()
    try:
        assert r.seam_steps(lib.oneCompleteStep, 60) == 0
    finally:
        lib.frehd_seam_release()
    dev = DeviceScenario(sc, from_state=init, precond=abi.PRECOND_SSOR)
    dev.step(60)
    assert_same(r, dev, sub, name)
    dev.close()


@pytest.mark.parametrize("name", list(CASES))
def test_solve_all_on_device(name):
    lib = seam()
    nt = 60
    sc = short_tide(CASES[name]())
    sub = sc.subgrid is not None
    print(f"case {name}: SolveAllOnDevice vs resident, {nt} steps", flush=True)
    r = po.Reference(sc, Nt=nt)
    init = ref_state(r, sub)
    assert r.seam_solve_all(lib.SolveAllOnDevice) == 0
    # the resident run: DataOutput's offset round trip at step 0, forcing
    # sample nt repeating sample nt - 1
    off, N = sc.cfg["offset"], sc.N
    init["surf"] = init["surf"].copy()
    init["surf"][:N] = (init["surf"][:N] - off) + off
    last = dataclasses.replace(sc)
    for k in ("tideP", "tideM", "inflow", "tidalPS", "tidalMS"):
        a = getattr(sc, k).copy()
        a[nt] = a[nt - 1]
        setattr(last, k, a)
    dev = DeviceScenario(last, from_state=init, Nt=nt, precond=abi.PRECOND_SSOR)
    dev.step(nt)
    assert_same(r, dev, sub, name)
    dev.close()
