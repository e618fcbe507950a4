"""The handle's place inside a time step and the calls it allows
(frehdc_amd/csrc/step_state.h), checked on the CPU: the header is host-only
C++, compiled here with g++ behind a tiny C wrapper.

A step is the five phases of frehd_dev_phase; phase 0 already advances the
step counter (SolveAll's tt += 1, nssolve.c:61), so after it the handle is
inside a step.  Whole steps, uploads, retiling and a new step index are
valid only at a step boundary; phases only forward; nothing but the batch
drives a handle that belongs to one."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

WRAP = r"""
#include "step_state.h"
extern "C" {
int ss_after(int phase) { return frehd::step_state_after(phase); }
int ss_check(int run, int batch, int call, int phase) {
    return frehd::step_call_check(run, batch != 0, (frehd::StepCall)call, phase);
}
int ss_new_depth(int run) { return frehd::step_state_new_depth(run); }
int ss_new_scalar(int run) { return frehd::step_state_new_scalar(run); }
}
"""

CALL_STEP, CALL_PHASE, CALL_UPLOAD, CALL_TILING, CALL_AUTOTUNE, CALL_SET_INDEX = range(6)
ESTATE, EINVAL = -6, -1


@pytest.fixture(scope="module")
def ss(tmp_path_factory):
    d = tmp_path_factory.mktemp("step_state")
    src, lib = d / "wrap.cpp", d / "libss.so"
    src.write_text(WRAP)
    subprocess.run(["g++", "-std=c++17", "-O1", "-shared", "-fPIC",
                    "-I", os.path.join(ROOT, "frehdc_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    "-o", str(lib), str(src)], check=True)
    return C.CDLL(str(lib))


def walk(ss, phases):
    run = 0
    for p in phases:
        assert ss.ss_check(run, 0, CALL_PHASE, p) == 0, (run, p)
        run = ss.ss_after(p)
    return run


def test_phase_zero_leaves_the_boundary(ss):
    # the bug this guards against: after phase 0 the handle looked as if it
    # sat at a step boundary, so a whole step or an upload was accepted
    run = walk(ss, [0])
    assert run == 1
    for call in (CALL_STEP, CALL_UPLOAD, CALL_TILING, CALL_AUTOTUNE, CALL_SET_INDEX):
        assert ss.ss_check(run, 0, call, 0) == ESTATE, call


def test_full_step_returns_to_the_boundary(ss):
    assert walk(ss, [0, 1, 2, 3, 4]) == 0
    for call in (CALL_STEP, CALL_UPLOAD, CALL_TILING, CALL_AUTOTUNE, CALL_SET_INDEX):
        assert ss.ss_check(0, 0, call, 0) == 0, call


@pytest.mark.parametrize("done,bad", [(1, 0), (2, 0), (2, 1), (3, 2), (4, 0), (4, 3)])
def test_phases_only_forward(ss, done, bad):
    run = walk(ss, list(range(done)))
    assert ss.ss_check(run, 0, CALL_PHASE, bad) == ESTATE
    for p in range(done, 5):
        assert ss.ss_check(run, 0, CALL_PHASE, p) == 0


def test_phase_may_start_mid_step(ss):
    # a handle made from a mid-step state runs the phase under examination
    assert walk(ss, [2, 3, 4]) == 0
    assert ss.ss_after(2) == 3 and ss.ss_new_depth(3) and not ss.ss_new_scalar(3)


@pytest.mark.parametrize("p", [-1, 5, 99])
def test_phase_out_of_range(ss, p):
    assert ss.ss_check(0, 0, CALL_PHASE, p) == EINVAL


def test_batched_handle_refuses_solo_calls(ss):
    for call in (CALL_STEP, CALL_PHASE, CALL_UPLOAD, CALL_TILING, CALL_AUTOTUNE, CALL_SET_INDEX):
        assert ss.ss_check(0, 1, call, 0) == ESTATE, call


def test_download_buffers_between_phases(ss):
    # updateData (phase 2) writes the new depth, scalarTransport (phase 3)
    # the new concentrations; a download in between reads those buffers
    expect = {0: (0, 0), 1: (0, 0), 2: (0, 0), 3: (1, 0), 4: (1, 1)}
    for done, (dep, sca) in expect.items():
        run = walk(ss, list(range(done)))
        assert (ss.ss_new_depth(run), ss.ss_new_scalar(run)) == (dep, sca), done
