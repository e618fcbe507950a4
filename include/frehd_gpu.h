/*
 * frehd_gpu.h -- C ABI of the MI355X (gfx950) semi-implicit free-surface step.
 *
 * The library replaces one Casulli time step of FrehdC v4.2, i.e. the body of
 * the time loop in FrehdC_v42/nssolve.c:61-104 (SolveAll):
 *
 *     oneCompleteStep(...)   nssolve.c:110-118  (solveSourceTerm, solveFreeSurface,
 *                                                updateData, scalarTransport)
 *     monitorCFL(...)        nsfunctions.c:1192-1232
 *
 * State lives on the device between calls and keeps the reference's flat
 * array layout (FrehdC_v42/map.c:15-128): interior cell ii = j*nx + i for
 * 0 <= ii < N = nx*ny, followed by the ghost blocks
 *     [N, N+nx)              jP ghosts (top row, j = ny)
 *     [N+nx, N+2nx)          jM ghosts (bottom row, j = -1)
 *     [N+2nx, N+2nx+ny)      iP ghosts (right column, i = nx)
 *     [N+2nx+ny, N+2nx+2ny)  iM ghosts (left column, i = -1)
 *     Ng-4 .. Ng-1           corners iMjP, iPjP, iPjM, iMjM
 * with Ng = (nx+2)(ny+2).  Face arrays (uuXP, depthXP, ...) hold the i+1/2
 * (X) or j+1/2 (Y) face of the cell with the same index.
 *
 * Every entry point returns 0 on success or a negative FREHD_E* code; the
 * message of the last failure on the calling thread is frehd_last_error().
 * Handles are not thread-safe: drive one handle from one host thread.
 */
#ifndef FREHD_GPU_H
#define FREHD_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes -------------------------------------------------------- */
#define FREHD_OK        0
#define FREHD_EINVAL   (-1)  /* bad argument or unsupported configuration      */
#define FREHD_EHIP     (-2)  /* a HIP runtime call failed                       */
#define FREHD_ENOMEM   (-3)  /* device or host allocation failed               */
#define FREHD_ECFL     (-4)  /* max CFL > 20 after a step; the reference prints
                                a warning and calls exit(EXIT_FAILURE) instead
                                (nsfunctions.c:1226-1231)                       */
#define FREHD_EBUDGET  (-5)  /* serving: request exceeds the HBM budget         */
#define FREHD_ESTATE   (-6)  /* call not valid in the handle's current state    */

/* ---- solver and precision selectors --------------------------------------- */
#define FREHD_PRECOND_JACOBI 0   /* laspack/precond.c:26-46 (JacobiPrecond)       */
#define FREHD_PRECOND_SSOR   1   /* laspack/precond.c:48-73 with omega = 1 (the
                                    reference's hard-wired choice, nsfunctions.c:1016):
                                    the triangular sweeps run diagonal by diagonal in
                                    one workgroup per scenario (nx + ny - 1 dependent
                                    stages), bit for bit the reference's vector;
                                    FP64 only                                        */

#define FREHD_PREC_FP64  0       /* every array and operation in FP64             */
#define FREHD_PREC_MIXED 1       /* FP32 stencils / PCG, FP64 refinement, FP64
                                    dot products, eta and volume accumulation;
                                    nx*ny < 2^29 (32-bit byte offsets into the
                                    FP32 vectors), else FREHD_EINVAL              */

/* ---- configuration ----------------------------------------------------------
 * Scalar fields of the reference Config (FrehdC_v42/configuration.h:4-111) that
 * the step reads, with the reference's names.  One strip only (np = 1), so
 * nx = NX and ny = NY.  Fields the reference never assigns (useEvap, useRain,
 * useSubStep, checkConservation, inflow_rank) are fixed at 0. */
typedef struct frehd_config {
    int32_t nx, ny;          /* interior cells in x and y                        */
    int32_t Nt;              /* steps covered by the forcing series               */
    int32_t maxIter;         /* PCG iteration cap                                 */
    double  dx, dy, dt;      /* cell size [m], time step [s]                      */
    double  g;               /* gravity                                           */
    double  NUx, NUy;        /* eddy viscosity                                    */
    double  CDx, CDy;        /* drag coefficient when CDnotN == 1                 */
    double  manningN;        /* Manning n when CDnotN == 0                        */
    int32_t CDnotN;          /* 0: Manning (updateCD), 1: constant CD             */
    int32_t useThinLayer;    /* thinLayerDrag each step (nsfunctions.c:691)       */
    double  hD, CDmax;       /* thin-layer height (Config.hD) and CD cap          */
    double  minDepth;        /* limiterCFL shallow-cell threshold                 */
    double  wtfh;            /* waterfall depth threshold                         */
    double  CFLl, CFLh;      /* advection CFL limiter bounds                      */
    double  eps;             /* relative residual stop: |r| < eps |b|             */
    int32_t precond;         /* FREHD_PRECOND_*                                   */
    int32_t precision;       /* FREHD_PREC_*                                      */
    int32_t bcType;          /* 1: tide on YP, 2: tide on YM, 3: both             */
    int32_t tideLocLengthP;  /* number of YP tidal cells                          */
    int32_t tideLocLengthM;  /* number of YM tidal cells                          */
    int32_t inflowLocLength; /* number of inflow cells                            */
    int32_t useScalar;       /* scalar transport on/off                           */
    int32_t scalarAdv;       /* 1: first-order upwind, 2: TVD superbee            */
    int32_t useConstTidePS, useConstTideMS;
    int32_t nScalar;         /* scalars carried (>= 1 when useScalar)             */
    double  tidePS, tideMS;  /* constant tidal scalar values                      */
    double  inflowS;         /* inflow scalar concentration                       */
    double  Kx, Ky;          /* scalar diffusivity                                */
    double  offset;          /* Bath.offset[0]: -min(bed) when the bed dips below 0 */
    /* refinement controls for FREHD_PREC_MIXED (ignored for FP64)               */
    double  epsInner;        /* FP32 inner PCG relative tolerance                 */
    int32_t maxRefine;       /* FP64 refinement sweeps cap                        */
    /* subgrid bathymetry (subgrid.c; Config.useSubgrid ... useminV)            */
    int32_t useSubgrid;      /* 0: plain grid, 1: subgrid tables                  */
    int32_t Ne;              /* table levels, round((surfmax - surfmin)/dsurf)    */
    int32_t warmStart;       /* FREHD_PREC_MIXED: refinement start; 0 the
                                reference's x = 0, 1 the current surface, 2 its
                                extrapolation 2 eta^n - eta^(n-1)                  */
    double  surfmin, dsurf;  /* level k is surfmin + k*dsurf (+ offset)           */
    double  subR;            /* dx / dxf (fine cells per coarse cell side)        */
    /* useSubgrid == 1 only: 0 keeps the Ne x N x 10 level tables on the device;
     * 1 keeps the fine bathymetry instead (frehd_subgrid.fine, fineR == subR)
     * and evaluates the two levels each lookup needs from the cell's pixels,
     * bitwise the values the tables hold.  For grids whose tables do not fit
     * (4096^2 at r = 4 and 450 levels: 600 GB of tables, 2 GB of pixels).
     * No counterpart in the reference, which reads tables from files. */
    int32_t subgridOnTheFly;
    /* wind stress (windTerm, nsfunctions.c:609-648; Config.useWind,
     * northAngle, rhoa, Cw, CwT): the series are frehd_forcing.windspd /
     * winddir; with useThinLayer the stress decays as exp(CwT (h - hD) / hD)
     * on faces shallower than hD */
    int32_t useWind;
    double  northAngle;      /* added to the wind direction [deg]                 */
    double  rhoa, Cw, CwT;   /* air density, wind drag coefficient, thin-layer rate */
    /* subgrid drag model (Config.useSubDrag, updateCDSub subgrid.c:1446-1535;
     * applies with useSubgrid = 1 and CDnotN = 0): 0 Manning with h = V/Z;
     * 2 Volp 2013, CD = g n^2 Yh capped at CDmax (frehd_subgrid.YhX/YhY);
     * 3, 4 Casas 2010, CD = g Cn^2 / (V/(dx dy))^(1/3) (frehd_subgrid.Cn).
     * Each takes its table's value at the lower bracketing level.  1 (the
     * curvature model) reads Sub.CvX/CvY, which the reference never fills,
     * and is rejected; the models need tables (subgridOnTheFly = 0). */
    int32_t useSubDrag;
    /* minimum face areas and effective volumes (Config.useminA, beta,
     * useminV; combineSubArea subgrid.c:360-414): useminA = 1 pulls a face
     * area above its level's Nmax/Omax back by beta; useminV = 1 reduces a
     * face volume where the minimum face area Nmin/Omin is positive, 2 takes
     * the minimum volumes.  They need the Nmin/Nmax/Omin/Omax tables
     * (useminA) and the minV* tables (useminA = 1 with useminV = 1, or
     * useminV = 2); with useminA = 0 and useminV = 0 or 1 nothing changes
     * (Nmin stays 0). */
    int32_t useminA, useminV;
    int32_t pad3;
    double  beta;
} frehd_config;

/* ---- forcing ----------------------------------------------------------------
 * Boundary series hold Nt+1 samples; sample tt is the forcing at step tt
 * (tNStart + tt*dt), the index the reference uses (nsfunctions.c:386, 921).
 * tideLocP/M and inflowLoc are interior cell indices (configuration.c:56-60,
 * 143-150).  manningCell is an optional per-cell Manning n (N values); NULL
 * means the uniform manningN. */
typedef struct frehd_forcing {
    const int32_t *tideLocP, *tideLocM, *inflowLoc;
    const double  *tideP, *tideM, *inflow;
    const double  *tidalPS, *tidalMS;   /* used when useConstTide?S == 0          */
    const double  *manningCell;
    const double  *windspd, *winddir;   /* useWind == 1: speed [m/s], direction
                                           from north [deg] (bc->windspd/winddir) */
} frehd_forcing;

/* ---- subgrid tables (initAllSubVar, subgrid.c:2112-2284) --------------------
 * Level-major tables: value of cell ii at level k is t[k*N + ii] (N interior
 * cells, Ne levels).  edgeX/edgeY are the raw (offset-free) bottom
 * elevations of the east and north faces (bottomXP.dat / bottomYP.dat).
 * staggeredV = 1 (the reference's only live form).  The minV* tables are
 * needed with useminA = 1 and useminV = 1, or useminV = 2; otherwise they
 * never change a result and may be NULL.  The drag-model and minimum-area
 * tables are needed only by the options that use them (frehd_config). */
typedef struct frehd_subgrid {
    const double *edgeX, *edgeY;
    const double *V, *Z, *Vxp, *Vxm, *Vyp, *Vym, *Np, *Nm, *Op, *Om;
    const double *minVxp, *minVxm, *minVyp, *minVym;
    /* Alternative to the ten tables (V ... Om all NULL): the fine bathymetry,
     * (ny*fineR) rows of (nx*fineR) raw elevations, x fastest, from which the
     * device generates the tables (the reference's preprocessing script,
     * script_CreateSubAreaWithBlockV16.m:129-262 with checkBlock = 0).  dxf
     * is the fine pixel size.  No counterpart in the reference's C code. */
    const double *fine;
    int32_t fineR, pad;
    double dxf;
    /* drag-model tables, level-major like the others (YhX.dat, YhY.dat, Cn.dat;
     * initAllSubVar subgrid.c:2206-2225): YhX/YhY with useSubDrag = 2, Cn with
     * 3 or 4, NULL otherwise */
    const double *YhX, *YhY, *Cn;
    /* minimum / maximum face-area tables (Nmin.dat ...; initAllSubVar
     * subgrid.c:2160-2190), level-major, with useminA = 1 */
    const double *Nmin, *Nmax, *Omin, *Omax;
} frehd_subgrid;

/* ---- initial condition (the inputs of Init, initialize.c:40-155) ------------
 * bottomZ is the raw bed elevation of the N interior cells as in bath.dat;
 * the offset is cfg->offset.  A per-cell initial field overrides the
 * constant when its pointer is not NULL.  S0cell holds nScalar*N values. */
typedef struct frehd_ic {
    const double *bottomZ;
    double surf0, U0, V0, S0;
    const double *surf0cell, *U0cell, *V0cell, *S0cell;
    const frehd_subgrid *sub;   /* required when cfg->useSubgrid == 1 */
} frehd_ic;

/* ---- state arrays (Data + Bath, initialize.h:10-20, bathymetry.h:6-14) -------
 * For frehd_dev_download every non-NULL pointer receives that array.  Arrays
 * marked Ng hold the full ghosted layout; the others hold N interior values.
 * S and Sm hold nScalar consecutive arrays. */
typedef struct frehd_fields {
    double *surf, *surfOld;            /* Ng */
    double *uuXP, *vvYP;               /* Ng */
    double *uuYP, *vvXP;               /* N  */
    double *depth;                     /* N  */
    double *depthXP, *depthYP;         /* Ng */
    double *EnXP, *EnYP;               /* Ng */
    double *DragXP, *DragYP;           /* Ng */
    double *CDXP, *CDYP;               /* Ng */
    double *Fuu, *Fvv;                 /* Ng */
    double *cellV;                     /* N  */
    double *S;                         /* nScalar x Ng */
    double *Sm;                        /* nScalar x N  */
    double *CFLx, *CFLy;               /* Ng (interior written) */
    double *wtfXP, *wtfYP;             /* Ng */
    double *bottomZ, *bottomXP, *bottomYP; /* Ng */
    /* subgrid state (Sub, subgrid.h:11-37), Ng each; ignored on a plain grid */
    double *V, *Z, *Vx, *Vy, *Zx, *Zy, *Nx, *Oy;
    int32_t *ind;                      /* Ng: level search start per cell     */
} frehd_fields;

/* ---- per-handle statistics -------------------------------------------------- */
typedef struct frehd_stats {
    int64_t steps;            /* steps taken since create                         */
    int64_t pcgIterTotal;     /* inner PCG iterations summed over all steps       */
    int32_t pcgIterLast;      /* inner iterations of the last step                */
    int32_t refineLast;       /* FP64 refinement sweeps of the last step (mixed)  */
    double  relResLast;       /* |r|/|b| at exit of the last solve                */
    double  maxCFLx, maxCFLy; /* from the last monitorCFL                         */
    double  Vloss[4];         /* Vloss1..4 accumulators (nsfunctions.c:446-486,1147) */
    int32_t halted;           /* 1 once max CFL > 20 stopped the handle            */
    int32_t haltStep;         /* the step at which it stopped                      */
} frehd_stats;

/* ---- conservation budgets (frehd_dev_set_budget / frehd_dev_budget) --------
 * Flux form, accumulated over the steps since the budgets were switched on:
 * per step, the change of the domain's total cell volume (volumeByFlux,
 * nsfunctions.c:373-409; subgrid.c:1227-1265) minus the flow through the
 * domain boundary's faces, the inflow and what the clip of negative cell
 * volumes added; per scalar, the change of its total mass over the
 * advection-diffusion update (scalar.c:68-274, with the inflow mass,
 * scalar.c:346-356) minus the advective and diffusive fluxes through the
 * boundary's faces and the inflow mass.  Interior faces cancel between
 * their cells, so the residuals measure only the rounding of the updates:
 * each step's is summed in double-double, the steps' in FP64.  The
 * concentration limiter and the out-of-range reset that follow
 * (scalar.c:357-412) are outside the flux form and outside these sums. */
#define FREHD_MAX_SCALARS 16
typedef struct frehd_budget {
    int64_t steps;              /* steps accounted                                */
    int32_t nScalar, on;
    double volume;              /* total cell volume after the last step [m^3]    */
    double volumeResidual;      /* sum of the steps' residuals [m^3]              */
    double volumeResidualMax;   /* largest |residual| of one step [m^3]           */
    double volumeBoundary;      /* volume in through the boundary faces [m^3]     */
    double volumeInflow;        /* inflow volume [m^3]                            */
    double volumeClip;          /* volume the clip of negative cells added [m^3]  */
    double volumeMoved;         /* sum over steps and cells of |volume change|:
                                   the water the faces moved [m^3]               */
    double mass[FREHD_MAX_SCALARS];            /* total Sm after the last step     */
    double massResidual[FREHD_MAX_SCALARS];
    double massResidualMax[FREHD_MAX_SCALARS];
    double massBoundary[FREHD_MAX_SCALARS];    /* mass in through the boundary     */
    double massInflow[FREHD_MAX_SCALARS];
    double massMoved[FREHD_MAX_SCALARS];       /* sum of |mass change|             */
} frehd_budget;

typedef struct frehd_dev frehd_dev_t;

/* Last error message on the calling thread ("" when none). */
const char *frehd_last_error(void);

/* Library version string. */
const char *frehd_version(void);

/* Number of visible devices (0 when none). */
int frehd_device_count(int32_t *count);

/* sizeof of frehd_config, frehd_forcing, frehd_ic, frehd_fields, frehd_stats,
 * frehd_budget as compiled into the library (binding checks). */
int frehd_abi_sizes(int64_t sizes[6]);

/* Create a device handle from raw inputs and run the Init sequence
 * (initialize.c:40-155: bathymetry ghosts, face elevations, initial fields,
 * boundary values, depths, volume, interpolated velocities, thin-layer drag,
 * scalar mass) on the device.  Replaces Init + the per-step LASpack objects
 * that SolveAll builds (nssolve.c:72-85). */
int frehd_dev_create(const frehd_config *cfg, const frehd_forcing *forcing,
                     const frehd_ic *ic, frehd_dev_t **out);

/* Create a handle from a state that the caller already initialised (the
 * reference's own Init output): every array of `state` must be non-NULL.
 * This is the drop-in form: state produced by FrehdC's Init is uploaded once.
 * With cfg->useSubgrid == 1, `sub` gives the tables and the subgrid state
 * arrays (V ... Oy, ind) of `state` are required; `sub` may be NULL otherwise. */
int frehd_dev_create_from_state(const frehd_config *cfg, const frehd_forcing *forcing,
                                const frehd_subgrid *sub, const frehd_fields *state,
                                frehd_dev_t **out);

/* Advance nsteps steps.  Each step is oneCompleteStep (nssolve.c:110-118)
 * followed by monitorCFL (nsfunctions.c:1192).  Step numbers continue from
 * the previous call (the first step is tt = 1).  Returns FREHD_ECFL, leaving
 * the state after the offending step, where the reference would exit. */
int frehd_dev_step(frehd_dev_t *dev, int32_t nsteps);

/* Run one phase of the next step and wait for it (diagnostics and parity
 * tests): 0 = solveSourceTerm (nssolve.c:121; tt advances here), 1 =
 * solveFreeSurface (:158), 2 = updateData (:181; also fills CFLx/CFLy),
 * 3 = scalarTransport (scalar.c:30), 4 = monitorCFL's stop test.  Running
 * 0..4 in order is one frehd_dev_step. */
int frehd_dev_phase(frehd_dev_t *dev, int32_t phase);

/* Copy the device state into the non-NULL arrays of `out` (DataOutput's
 * view, fileio.c:62-180, at full precision). */
int frehd_dev_download(frehd_dev_t *dev, frehd_fields *out);

/* Statistics of the handle. */
int frehd_dev_stats(frehd_dev_t *dev, frehd_stats *out);

/* Switch the conservation budgets on (on != 0) or off and clear them: the
 * steps after it are accounted (frehd_budget).  No counterpart in the
 * reference (its checkConservation flag is never assigned, configuration.c).
 * At a step boundary, on a handle not in a batch (set it before adding the
 * handle); costs a few per-block partial sums per step. */
int frehd_dev_set_budget(frehd_dev_t *dev, int32_t on);
int frehd_dev_budget(frehd_dev_t *dev, frehd_budget *out);

/* Current step counter (number of steps taken). */
int frehd_dev_step_index(frehd_dev_t *dev, int32_t *tt);

/* Replace state arrays of the handle with host arrays: every non-NULL
 * pointer of *in (the frehd_fields layout of frehd_dev_download) is copied
 * into the buffers the next step starts from, as frehd_dev_create_from_state
 * does at creation.  For callers that change the state on the host between
 * steps -- oneCompleteStep with the reference's signature (frehd_seam.h;
 * nssolve.c:110-118), whose caller runs monitorCFL and DataOutput on the
 * host.  The subgrid lookup then starts afresh for every cell.  At a step
 * boundary only (FREHD_ESTATE after a partial step, or for a handle in a
 * batch). */
int frehd_dev_upload(frehd_dev_t *dev, const frehd_fields *in);

/* Whole steps as frehd_dev_step, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (serving loops
 * and harnesses that must not call hipMalloc).  The kernels' per-block
 * partial sums go to a static device array of the library instead of the
 * handle's buffers -- 16 MiB, enough for grids up to about 4096 x 4096 cells
 * (FREHD_EINVAL beyond) -- and the kernels are launched one at a time (no
 * graph is instantiated), untimed.  Results are bit for bit those of
 * frehd_dev_step; the two may be mixed on one handle.  The array is shared
 * by every call on a device and keeps the last call's contents (each call
 * clears what it uses first): calls that overlap on one device race.
 * FREHD_ESTATE mid-step or for a handle in a batch. */
int frehd_dev_step_noalloc(frehd_dev_t *dev, int32_t nsteps);

/* Set the step counter, e.g. to resume from a state saved at step tt; the
 * next step taken is tt + 1. */
int frehd_dev_set_step_index(frehd_dev_t *dev, int32_t tt);

/* Release the handle and its device memory. */
int frehd_dev_destroy(frehd_dev_t *dev);

/* Device bytes a handle of this configuration occupies. */
int frehd_dev_footprint(const frehd_config *cfg, int64_t *bytes);

/* Mean device time per launch of each kernel of the last frehd_dev_step call,
 * measured with HIP events on the handle's stream when timing is enabled.
 * names/ms/launches receive up to `cap` entries; *n gets the count. */
int frehd_dev_enable_timing(frehd_dev_t *dev, int32_t on);
int frehd_dev_kernel_times(frehd_dev_t *dev, int32_t cap, const char **names,
                           double *ms, int64_t *launches, int32_t *n);

/* Download the subgrid level tables of a handle (e.g. the ones generated on
 * the device): tables[0..9] receive V Z Vxp Vxm Vyp Vym Np Nm Op Om, Ne*N
 * values each, level-major; NULL entries are skipped.  No counterpart in the
 * reference (its tables come from files, subgrid.c:2112-2284). */
int frehd_dev_download_tables(frehd_dev_t *dev, double *const tables[10]);

/* Launch shape of the solver's direction kernel (no counterpart in the
 * reference; a tuning knob for the autotuner): tiles of 256 x cols_per_lane
 * columns x rows_per_tile rows (0 = automatic: about 1024 tiles, 512 for
 * the mixed path's inner solve), and xcd_remap != 0 to give each of the 8
 * XCDs a contiguous band of tiles.  cols_per_lane (1, 2 or 4: adjacent
 * columns each lane loads as one vector) applies to the mixed path's inner
 * FP32 solve and needs nx to be a multiple of it; the FP64 PCG takes 1.
 * Results do not depend on the shape beyond the order of the solver's
 * dot-product sums.  A handle joining a batch takes the members'
 * cols_per_lane (or, if its nx does not allow it, the batch goes back to 1). */
int frehd_dev_set_pcg_tiling(frehd_dev_t *dev, int32_t rows_per_tile, int32_t xcd_remap, int32_t cols_per_lane);

/* Autotuner of that launch shape (no counterpart in the reference): times
 * `iters` solver iterations (the FP64 PCG pair, or the mixed path's inner
 * FP32 iteration plus its share of FP64 refinement sweeps, in the
 * proportion of the last step's solve) under every candidate shape --
 * rows_per_tile 0, 1, 2, 4, ... up to min(ny, 256), XCD remap off and on,
 * cols_per_lane 1, 2, 4 (mixed path, where nx allows) -- on the system of
 * the last step, applies the fastest and returns it (each output may be
 * NULL; rows_per_tile as resolved).  The model state, the step counter and
 * the statistics are left as they were.  Call it after at least one step;
 * FREHD_ESTATE for a handle in a batch (tune the batch). */
int frehd_dev_autotune(frehd_dev_t *dev, int32_t iters, int32_t *rows_per_tile, int32_t *xcd_remap,
                       int32_t *cols_per_lane, double *ms_per_iteration);

/* Fused solve of the mixed-precision form (no counterpart in the reference):
 * the whole free-surface solve of a step -- the FP64 refinement sweeps, the
 * inner FP32 iterations, their convergence tests, the surface update and
 * the solver flags -- runs as one persistent launch with grid-wide barriers
 * instead of one launch per sweep or iteration; results are bitwise those of
 * the launched solve.  Its blocks must be resident on the device together:
 * max_blocks < 0 switches it off (the launched solve, where a handle
 * starts: on MI355X a grid-wide barrier costs more than the kernel boundary
 * it replaces), 0 lets it take as many blocks as the device holds at once,
 * > 0 caps it.  A
 * caller that steps several handles or batches concurrently (one stream
 * each) divides frehd_fused_solve_capacity among them; a launch whose blocks
 * were not resident together ends after a 2 s barrier wait with FREHD_EHIP.
 * The launched solve's pattern history (launch counts of the last solves)
 * starts afresh.
 * No effect on FP64 handles; FREHD_ESTATE for a handle in a batch (set it on
 * the batch: frehd_batch_set_fused_solve). */
int frehd_dev_set_fused_solve(frehd_dev_t *dev, int32_t max_blocks);
/* blocks of the fused solve the current device holds at once, for a solve
 * with cols_per_lane (1, 2, 4) columns per lane */
int frehd_fused_solve_capacity(int32_t cols_per_lane, int32_t *blocks);

/* Block shape of the 2-D stencil kernels (predictor, limiter, scalar
 * transport; no counterpart in the reference): each candidate -- blocks of
 * 64 x 4, 128 x 2 and 256 x 1 cells -- runs `steps` whole steps twice with
 * per-kernel event timing, and the shape whose three kernels take the least
 * time per step is kept (block_x: 64, 128 or 256; ms_per_step: their time;
 * either output may be NULL).  The model ADVANCES by 6 * steps steps: the
 * shape changes no cell's arithmetic, only the order of the block partial
 * sums of the volume losses and budgets, so these are ordinary steps of the
 * run.  FREHD_EINVAL when the forcing series ends before them; FREHD_ESTATE
 * for a handle in a batch (tune the batch). */
int frehd_dev_autotune_blocks(frehd_dev_t *dev, int32_t steps, int32_t *block_x, double *ms_per_step);

/* ---- batches: many handles advanced by shared launches ---------------------
 * No counterpart in the reference (one scenario per process there).  A batch
 * holds up to `capacity` handles of one device that share useSubgrid and
 * precision (grid sizes, forcing and physics may differ); every kernel of a
 * step is launched once for all of them.  Handles stay usable on their own
 * (download, stats, ...) between batch calls; a handle is in at most one
 * batch at a time, and removing it (or destroying the batch) leaves it
 * intact.  frehd_batch_step advances every member that is not halted by
 * nsteps steps and returns FREHD_ECFL if a member is halted at the end (its
 * frehd_dev_stats / frehd_dev_step_index tell which); FREHD_EINVAL if a
 * member's forcing series is too short for nsteps, before any step. */
typedef struct frehd_batch frehd_batch_t;

int frehd_batch_create(int32_t capacity, frehd_batch_t **out);
int frehd_batch_add(frehd_batch_t *batch, frehd_dev_t *dev);
int frehd_batch_remove(frehd_batch_t *batch, frehd_dev_t *dev);
int frehd_batch_size(frehd_batch_t *batch, int32_t *n);
int frehd_batch_step(frehd_batch_t *batch, int32_t nsteps);
/* frehd_dev_autotune over the members together (one shape for all: the
 * launches are shared). */
int frehd_batch_autotune(frehd_batch_t *batch, int32_t iters, int32_t *rows_per_tile, int32_t *xcd_remap,
                         int32_t *cols_per_lane, double *ms_per_iteration);
/* frehd_dev_autotune_blocks for a batch (one shape for all members; the
 * members advance by 6 * steps steps) */
int frehd_batch_autotune_blocks(frehd_batch_t *batch, int32_t steps, int32_t *block_x, double *ms_per_step);
/* frehd_dev_set_fused_solve for a batch (one fused launch for all members) */
int frehd_batch_set_fused_solve(frehd_batch_t *batch, int32_t max_blocks);
int frehd_batch_enable_timing(frehd_batch_t *batch, int32_t on);
int frehd_batch_kernel_times(frehd_batch_t *batch, int32_t cap, const char **names,
                             double *ms, int64_t *launches, int32_t *n);
int frehd_batch_destroy(frehd_batch_t *batch);

#ifdef __cplusplus
}
#endif

#endif /* FREHD_GPU_H */
