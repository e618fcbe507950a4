// step_state.h -- where a handle stands inside a time step, and which entry
// points that allows (host only; no HIP).
//
// A step is the five phases of frehd_dev_phase, 0..4 (solveSourceTerm,
// solveFreeSurface, updateData, scalarTransport, monitorCFL's stop test).
// `phasesRun` counts the phases of the step in progress: 0 at a step
// boundary, p + 1 once phase p (p < 4) has run; phase 4 closes the step and
// brings it back to 0.  Phase 0 already advances the step counter, so a
// handle with phasesRun > 0 is inside a step even though its counter moved.
#pragma once

#include "frehd_gpu.h"

namespace frehd {

enum StepCall {
    CALL_STEP,      // frehd_dev_step
    CALL_PHASE,     // frehd_dev_phase(phase)
    CALL_UPLOAD,    // frehd_dev_upload
    CALL_TILING,    // frehd_dev_set_pcg_tiling
    CALL_AUTOTUNE,  // frehd_dev_autotune
    CALL_SET_INDEX  // frehd_dev_set_step_index
};

// phasesRun after phase `phase` ran
inline int step_state_after(int phase) { return phase >= 4 ? 0 : phase + 1; }

// 0 when `call` is valid for a handle with `phasesRun` phases of its step
// done, FREHD_ESTATE when it is not, FREHD_EINVAL for a phase outside 0..4.
// A handle in a batch is driven by the batch.  Whole steps, uploads and
// retiling only at a step boundary.  Phases only forward: none of the step
// in progress again (phase 0 again would start another step and skip a
// forcing sample), while a later one may follow directly -- a parity check
// starts a handle created from a mid-step state at the phase it examines.
inline int step_call_check(int phasesRun, bool inBatch, StepCall call, int phase = 0)
{
    if (call == CALL_PHASE && (phase < 0 || phase > 4)) return FREHD_EINVAL;
    if (inBatch) return FREHD_ESTATE;
    if (call == CALL_PHASE) return phase >= phasesRun ? 0 : FREHD_ESTATE;
    return phasesRun == 0 ? 0 : FREHD_ESTATE;
}

// which double buffers a download reads between phases: the surface of
// solveFreeSurface (phase 1), the depth updated by updateData (phase 2) and
// the concentrations of scalarTransport (phase 3)
inline bool step_state_new_surface(int phasesRun) { return phasesRun >= 2; }
inline bool step_state_new_depth(int phasesRun) { return phasesRun >= 3; }
inline bool step_state_new_scalar(int phasesRun) { return phasesRun >= 4; }

}  // namespace frehd
