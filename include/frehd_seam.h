/*
 * frehd_seam.h -- FrehdC v4.2's own time-step entry points, served by the
 * MI355X device library (frehd_gpu.h).  A maintainer compiles
 * frehdc_amd/seam/frehd_seam.c inside the FrehdC tree, against its headers
 * (Config, Data, Sub, Maps, BC, Bath; LASpack's QMatrix and Vector), and
 * drops nssolve.c's definitions of the two functions below.
 *
 *   oneCompleteStep   nssolve.h:10-11 / nssolve.c:110-118, the reference's
 *                     signature exactly (declared by nssolve.h itself):
 *       void oneCompleteStep(Data **data, Sub **sub, Maps *map, BC *bc,
 *                            Bath *bath, Config *setting, int tt, QMatrix A,
 *                            Vector x, Vector z, int irank, int nrank);
 *                     The host state goes to the device, one step (the four
 *                     phases) runs there, the state comes back; A, x and z
 *                     are not used (the device's solve is matrix-free).  The
 *                     caller's monitorCFL and DataOutput run on the host as
 *                     in SolveAll (nssolve.c:82-95).
 *   SolveAllOnDevice  SolveAll (nssolve.h:9, nssolve.c:48-106) with the
 *                     state resident on the device between outputs: the
 *                     efficient seam (no state round trip per step).
 *   frehd_seam_release  no counterpart in the reference: frees the device
 *                     handle oneCompleteStep keeps between calls (the next
 *                     call starts a new one from the host state).
 *
 * One strip (np = 1: irank = 0, nrank = 1), one scalar, the reference's
 * Manning n, and the reference's solve: CG with SSOR (omega = 1,
 * nsfunctions.c:1016).  The environment variable FREHD_SEAM_PRECOND=jacobi
 * selects Jacobi-preconditioned CG instead (same tolerance; the device's SSOR
 * sweeps are nx + ny - 1 dependent stages, slow on large grids).
 */
#ifndef FREHD_SEAM_H
#define FREHD_SEAM_H

#include "configuration.h"
#include "bathymetry.h"
#include "initialize.h"
#include "map.h"
#include "subgrid.h"
#include "nssolve.h"

void SolveAllOnDevice(Data **data, Sub **sub, Maps *map, BC *bc, Bath *bath, Config *setting, int irank, int nrank);
void frehd_seam_release(void);

#endif /* FREHD_SEAM_H */
