"""The oneCompleteStep shim (the reference's signature, nssolve.c:110-118,
forwarding to the device library): per step the host state goes to the
device (frehd_dev_upload), the four phases of oneCompleteStep and the
device's monitorCFL with the step's end (phase 4) run there, the state comes
back, and the host runs monitorCFL (nsfunctions.c:1192-1232, CFL = |u| dt/dx
per face) as the reference's loop does, before the next call.  Over a wet/dry run this must
give the resident path's results (frehd_dev_step) bit for bit: the same
kernels on the same inputs."""
import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import DeviceScenario

pytestmark = pytest.mark.gpu

STATE = ["surf", "surfOld", "uuXP", "vvYP", "uuYP", "vvXP", "depth", "depthXP", "depthYP", "EnXP", "EnYP",
         "DragXP", "DragYP", "CDXP", "CDYP", "Fuu", "Fvv", "cellV", "S", "Sm", "CFLx", "CFLy", "wtfXP", "wtfYP"]


@pytest.mark.parametrize("make", [lambda: cases.marsh(n=40, dx=2.0, dt=0.5, Nt=120, amp=0.6),
                                  lambda: cases.marsh_subgrid(n=24, r=4, Nt=120, period=1200.0)],
                         ids=["marsh40", "marsh24_sub4"])
def test_one_complete_step_shim(make):
    sc = make()
    print(f"case {sc.name}: shim vs resident, 60 steps", flush=True)
    names = STATE + (["V", "Z", "Vx", "Vy", "Zx", "Zy", "Nx", "Oy", "ind"] if sc.subgrid is not None else [])
    res = DeviceScenario(sc)
    shim = DeviceScenario(sc)
    host = shim.download(names)
    N, Ng = sc.N, sc.Ng
    dt, dx, dy = sc.cfg["dt"], sc.cfg["dx"], sc.cfg["dy"]
    for t in range(60):
        shim.upload(host)
        for p in range(5):                 # oneCompleteStep (phases 0..3) + monitorCFL
            shim.phase(p)
        host = shim.download(names)
        # monitorCFL on the host (nsfunctions.c:1197-1206)
        host["CFLx"][:N] = np.abs(host["uuXP"][:N]) * dt / dx
        host["CFLy"][:N] = np.abs(host["vvYP"][:N]) * dt / dy
    res.step(60)
    got = res.download(names)
    for n in names:
        a, b = np.asarray(got[n]).reshape(-1), np.asarray(host[n]).reshape(-1)
        if a.dtype == np.float64:
            assert np.array_equal(a.view(np.int64), b.view(np.int64)), n
        else:
            assert np.array_equal(a, b), n
    assert res.stats()["steps"] == 60
    res.close()
    shim.close()
