/*
 * ref_harness.c -- drives FrehdC v4.2's own step functions on given inputs.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
 * the reference's unmodified sources (FrehdC_v42/*.c except FREHD.c, plus the
 * LASpack files they use) from where they lie under the reference checkout,
 * into oracle/_ref/libfrehd_ref.so.  Nothing of the reference is copied into
 * this repository; the built library stays out of git (.gitignore: *.so).
 *
 * The harness replaces only what cannot run as shipped (the reference's
 * main in FREHD.c:66-118 and Init in initialize.c:40-155):
 *   - Config is filled directly instead of parsed from data.dat
 *     (ReadUserSettings/dateNum read an unterminated buffer, utilities.c:114);
 *     the never-assigned flags useEvap, useRain, useSubStep,
 *     checkConservation and inflow_rank are set to 0;
 *   - the grid maps are built here with the same index formulas as
 *     createMaps (map.c:15-128) but with icjMM sized correctly (the reference
 *     writes past its end, map.c:69-74) and defined on row 1, which those
 *     loops never set: two cells down there is the jM ghost, as iMMjc is the
 *     iM ghost at i = 1;
 *   - every Data/Bath/BC array is allocated with zeroed guard cells before and
 *     after it, so the reference's out-of-range reads (velocityInterp's
 *     uuXP[-1] and vvYP[ii-nx]; limiterCFL's depth[ghost]) read 0.0;
 *   - boundary series are given per step (Nt+1 samples) instead of being read
 *     and interpolated from files;
 *   - the Init call sequence is restated, calling the reference's functions
 *     for every numerical part.
 * The ref_io_* functions call the reference's own input and output routines
 * (readOneBC, ReadBathymetry + InitBathymetry, initAllSubVar, WriteFile) on
 * given files, for the parity tests of the deck reader; paths must stay
 * short (readOneBC builds them in a 50-byte buffer, initialize.c:405).
 * The step itself is the reference's solveSourceTerm / solveFreeSurface /
 * updateData / scalarTransport / monitorCFL.  With the Jacobi preconditioner
 * the harness restates solveFreeSurface (nssolve.c:158-178) with
 * CGIter(..., JacobiPrecond, 1) in place of the hard-wired SSORPrecond
 * (nsfunctions.c:1016); everything else in that phase is the reference's.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "configuration.h"
#include "bathymetry.h"
#include "initialize.h"
#include "map.h"
#include "subgrid.h"
#include "nsfunctions.h"
#include "nssolve.h"
#include "scalar.h"
#include "laspack/errhandl.h"
#include "laspack/vector.h"
#include "laspack/qmatrix.h"
#include "laspack/operats.h"
#include "laspack/precond.h"
#include "laspack/rtc.h"
#include "laspack/itersolv.h"

#include "../include/frehd_gpu.h"

typedef struct ref_handle {
    Config *setting;
    Maps *map;
    Data *data;
    Bath *bath;
    BC *bc;
    Sub *sub;
    IC *ic;
    int tt, precond, lastIter;
    double *blocks[256];
    int nblocks;
} ref_handle;

/* zero-padded allocation: `pad` zero doubles before and after n values */
static double *padded(ref_handle *h, size_t n, size_t pad)
{
    double *b = (double *)calloc(n + 2 * pad, sizeof(double));
    h->blocks[h->nblocks++] = b;
    return b + pad;
}

static int *ialloc(size_t n) { return (int *)calloc(n, sizeof(int)); }

static void build_maps(Maps *m, int nx, int ny)
{
    int N = nx * ny, Ng = (nx + 2) * (ny + 2);
    m->cntr = ialloc(N); m->trps = ialloc(N); m->sprt = ialloc(N);
    m->iPjc = ialloc(N); m->iMjc = ialloc(N); m->icjP = ialloc(N); m->icjM = ialloc(N);
    m->iPPjc = ialloc(N); m->iMMjc = ialloc(N); m->icjPP = ialloc(N); m->icjMM = ialloc(N + nx);
    m->iPjP = ialloc(1); m->iPjM = ialloc(1); m->iMjP = ialloc(1); m->iMjM = ialloc(1);
    m->ii2d = ialloc(N); m->jj2d = ialloc(N);
    m->iPbd = ialloc(ny); m->iPgt = ialloc(ny); m->iMbd = ialloc(ny); m->iMgt = ialloc(ny);
    m->jPbd = ialloc(nx); m->jPgt = ialloc(nx); m->jMbd = ialloc(nx); m->jMgt = ialloc(nx);
    for (int c = 0; c < N; c++) {
        int i = c % nx, j = c / nx;
        m->cntr[c] = c;
        m->trps[c] = (c % ny) * nx + c / ny;
        m->sprt[c] = (c % nx) * ny + c / nx;
        m->icjP[c] = j == ny - 1 ? N + i : c + nx;
        m->icjM[c] = j == 0 ? N + nx + i : c - nx;
        m->iPjc[c] = i == nx - 1 ? N + 2 * nx + j : c + 1;
        m->iMjc[c] = i == 0 ? N + 2 * nx + ny + j : c - 1;
        m->ii2d[c] = i; m->jj2d[c] = j;
    }
    for (int c = 0; c < N; c++) {
        int i = c % nx, j = c / nx;
        m->icjPP[c] = j >= ny - 1 ? m->icjP[c] : m->icjP[m->icjP[c]];
        m->icjMM[c] = j <= 0 ? m->icjM[c] : m->icjM[m->icjM[c]];
        m->iPPjc[c] = i == nx - 1 ? m->iPjc[c] : m->iPjc[m->iPjc[c]];
        m->iMMjc[c] = i == 0 ? m->iMjc[c] : m->iMjc[m->iMjc[c]];
    }
    m->iMjM[0] = Ng - 1; m->iPjM[0] = Ng - 2; m->iPjP[0] = Ng - 3; m->iMjP[0] = Ng - 4;
    for (int i = 0; i < nx; i++) {
        m->jPbd[i] = N - nx + i; m->jMbd[i] = i;
        m->jPgt[i] = m->icjP[N - nx + i]; m->jMgt[i] = m->icjM[i];
    }
    for (int j = 0; j < ny; j++) {
        m->iPbd[j] = (j + 1) * nx - 1; m->iMbd[j] = j * nx;
        m->iPgt[j] = m->iPjc[(j + 1) * nx - 1]; m->iMgt[j] = m->iMjc[j * nx];
    }
}

static void fill_config(Config *s, const frehd_config *c)
{
    memset(s, 0, sizeof(*s));
    s->NX = s->nx = c->nx; s->NY = s->ny = c->ny;
    s->N2CI = s->N2ci = c->nx * c->ny;
    s->N2CT = s->N2ct = (c->nx + 2) * (c->ny + 2);
    s->dx = c->dx; s->dy = c->dy; s->dt = c->dt; s->Nt = c->Nt; s->OutItvl = c->Nt + 1;
    s->g = c->g; s->NUx = c->NUx; s->NUy = c->NUy; s->CDx = c->CDx; s->CDy = c->CDy;
    s->manningN = c->manningN; s->CDnotN = c->CDnotN; s->hD = c->hD; s->CDmax = c->CDmax;
    s->useThinLayer = c->useThinLayer; s->minDepth = c->minDepth; s->wtfh = c->wtfh;
    s->CFLl = c->CFLl; s->CFLh = c->CFLh; s->eps = c->eps; s->maxIter = c->maxIter;
    s->bcType = c->bcType; s->tideLocLengthP = c->tideLocLengthP;
    s->tideLocLengthM = c->tideLocLengthM; s->inflowLocLength = c->inflowLocLength;
    s->useScalar = c->useScalar; s->scalarAdv = c->scalarAdv;
    s->useConstTidePS = c->useConstTidePS; s->useConstTideMS = c->useConstTideMS;
    s->tidePS = c->tidePS; s->tideMS = c->tideMS; s->inflowS = c->inflowS;
    s->Kx = c->Kx; s->Ky = c->Ky;
    s->useSubgrid = c->useSubgrid; s->useMPI = 0; s->np = 1; s->useWind = c->useWind;
    s->northAngle = c->northAngle; s->rhoa = c->rhoa; s->Cw = c->Cw; s->CwT = c->CwT;
    s->staggeredV = 1; s->useminA = c->useminA; s->useminV = c->useminV; s->useSubDrag = c->useSubDrag;
    s->beta = c->beta;
    s->dsurf = c->dsurf; s->surfmin = c->surfmin; s->surfmax = c->surfmin + c->Ne * c->dsurf;
    s->subR = c->subR; s->dxf = c->dx / c->subR; s->dyf = c->dy / c->subR;
    s->dA = s->dxf * s->dyf; s->z0 = c->hD;
    s->useEvap = 0; s->useRain = 0; s->useSubStep = 0; s->checkConservation = 0;
    s->inflow_rank = 0; s->useCellEdge = 0;
}

void ref_destroy(void *hv);

void *ref_create(const frehd_config *c, const frehd_forcing *f, const frehd_ic *icin)
{
    if (c->tideLocLengthP > 1000 || c->tideLocLengthM > 1000 || c->inflowLocLength > 1000)
        return NULL;    /* Config holds at most 1000 locations (configuration.h:60,68) */
    if (c->useScalar && c->nScalar > 1) return NULL;
    ref_handle *h = (ref_handle *)calloc(1, sizeof(ref_handle));
    h->precond = c->precond;
    const int nx = c->nx, ny = c->ny, N = nx * ny, Ng = (nx + 2) * (ny + 2);
    const size_t pad = (size_t)nx + 8;
    h->setting = (Config *)calloc(1, sizeof(Config));
    fill_config(h->setting, c);
    Config *s = h->setting;
    for (int q = 0; q < c->tideLocLengthP; q++) s->tideLocP[q] = f->tideLocP[q];
    for (int q = 0; q < c->tideLocLengthM; q++) s->tideLocM[q] = f->tideLocM[q];
    for (int q = 0; q < c->inflowLocLength; q++) s->inflowLoc[q] = f->inflowLoc[q];

    h->map = (Maps *)calloc(1, sizeof(Maps));
    build_maps(h->map, nx, ny);

    /* Bath (bathymetry.c:20-104) */
    h->bath = (Bath *)calloc(1, sizeof(Bath));
    Bath *b = h->bath;
    b->offset = padded(h, 1, 1);
    b->offset[0] = c->offset;
    b->allZ = padded(h, N, pad);
    b->bottomZ = padded(h, Ng, pad);
    b->bottomXP = padded(h, Ng, pad);
    b->bottomYP = padded(h, Ng, pad);
    b->edgeX = padded(h, Ng, pad);
    b->edgeY = padded(h, Ng, pad);
    for (int q = 0; q < N; q++) b->allZ[q] = icin->bottomZ[q];
    for (int q = 0; q < N; q++) b->bottomZ[q] = b->allZ[q] + b->offset[0];
    b->allEdgeX = padded(h, N, pad);
    b->allEdgeY = padded(h, N, pad);
    if (c->useSubgrid) {
        if (!icin->sub || !icin->sub->edgeX || !icin->sub->edgeY) return NULL;
        for (int q = 0; q < N; q++) {
            b->allEdgeX[q] = icin->sub->edgeX[q];
            b->allEdgeY[q] = icin->sub->edgeY[q];
            b->edgeX[q] = b->allEdgeX[q] + b->offset[0];
            b->edgeY[q] = b->allEdgeY[q] + b->offset[0];
        }
    }

    /* Data arrays (initDataArrays, initialize.c:178-270), zero padded */
    h->data = (Data *)calloc(1, sizeof(Data));
    Data *d = h->data;
    double **arrs[] = {&d->uuXP, &d->uuYP, &d->vvXP, &d->vvYP, &d->surf, &d->surfOld, &d->Fuu,
        &d->Fvv, &d->cellV, &d->depth, &d->depthXP, &d->depthYP, &d->EnXP, &d->EnYP,
        &d->DragXP, &d->DragYP, &d->CDXP, &d->CDYP, &d->GnCt, &d->GnXP, &d->GnYP, &d->GnXM,
        &d->GnYM, &d->z, &d->wtfXP, &d->wtfYP, &d->S, &d->Sm, &d->rmvCFL, &d->advX, &d->advY,
        &d->CFLx, &d->CFLy, &d->Vloss, &d->alluuXP, &d->allvvYP, &d->allsurf, &d->alldepth,
        &d->allS, &d->allCDXP, &d->allCDYP, &d->allVloss};
    for (size_t q = 0; q < sizeof(arrs) / sizeof(arrs[0]); q++) *arrs[q] = padded(h, Ng, pad);
    d->Vloss1 = padded(h, 1, 1); d->Vloss2 = padded(h, 1, 1);
    d->Vloss3 = padded(h, 1, 1); d->Vloss4 = padded(h, 1, 1);
    d->inflowLoc = ialloc(c->inflowLocLength + 1);

    /* BC series (initBCArrays, initialize.c:330-342), Nt+1 samples */
    h->bc = (BC *)calloc(1, sizeof(BC));
    BC *bc = h->bc;
    const size_t ns = (size_t)c->Nt + 1;
    double **bcs[] = {&bc->tideP, &bc->tideM, &bc->inflow, &bc->windspd, &bc->winddir,
        &bc->tidalPS, &bc->tidalMS, &bc->evap, &bc->rain};
    for (size_t q = 0; q < sizeof(bcs) / sizeof(bcs[0]); q++) *bcs[q] = padded(h, ns, 4);
    for (size_t t = 0; t < ns; t++) {
        if (f->tideP) bc->tideP[t] = f->tideP[t];
        if (f->tideM) bc->tideM[t] = f->tideM[t];
        if (f->inflow) bc->inflow[t] = f->inflow[t];
        if (f->tidalPS) bc->tidalPS[t] = f->tidalPS[t];
        if (f->tidalMS) bc->tidalMS[t] = f->tidalMS[t];
        if (c->useWind == 1 && f->windspd) bc->windspd[t] = f->windspd[t];
        if (c->useWind == 1 && f->winddir) bc->winddir[t] = f->winddir[t];
    }

    h->sub = (Sub *)calloc(1, sizeof(Sub));

    /* IC (readIC, initialize.c:279-327) */
    h->ic = (IC *)calloc(1, sizeof(IC));
    IC *ic = h->ic;
    s->useConstSurf0 = icin->surf0cell ? 0 : 1;
    s->useConstU0 = icin->U0cell ? 0 : 1;
    s->useConstV0 = icin->V0cell ? 0 : 1;
    s->useConstInitS = icin->S0cell ? 0 : 1;
    ic->surf0 = icin->surf0; ic->U0 = icin->U0; ic->V0 = icin->V0; ic->S0 = icin->S0;
    ic->allSurf0 = (double *)icin->surf0cell; ic->allU0 = (double *)icin->U0cell;
    ic->allV0 = (double *)icin->V0cell; ic->allS0 = (double *)icin->S0cell;

    /* Init's call sequence, initialize.c:40-155 (np = 1, plain grid) */
    enforceBathBC(b, h->map, s, 0, 1);
    updateAllFaceElev(b, h->map, s);
    initCD(&h->data, s);
    initFieldValues(&h->data, b, ic, s, 0);
    zeroNegSurf(&h->data, b, s, 0);
    enforceFreeSurfaceBC(&h->data, h->map, s);
    enforceVelocityBC(&h->data, h->map, s);
    enforceTidalBC(&h->data, b, bc, s, 0, 0, 1);
    updateDepth(&h->data, h->map, b, s);
    updateAllFaceDepth(&h->data, b, h->map, s, 0, 1);
    if (c->useSubgrid) {
        /* initAllSubVar (subgrid.c:2112-2284) without the file reads */
        Sub *sb = h->sub;
        const int Ne = c->Ne;
        sb->Ne = Ne;
        sb->allSurf = padded(h, Ne, 4);
        for (int q = 0; q < Ne; q++) sb->allSurf[q] = s->surfmin + q * s->dsurf + b->offset[0];
        const frehd_subgrid *sg = icin->sub;
        double **tabs[] = {&sb->allV, &sb->allZ, &sb->allVxp, &sb->allVxm, &sb->allVyp, &sb->allVym,
            &sb->allNp, &sb->allNm, &sb->allOp, &sb->allOm, &sb->allminVxp, &sb->allminVxm,
            &sb->allminVyp, &sb->allminVym};
        const double *src[] = {sg->V, sg->Z, sg->Vxp, sg->Vxm, sg->Vyp, sg->Vym, sg->Np, sg->Nm,
            sg->Op, sg->Om, sg->minVxp, sg->minVxm, sg->minVyp, sg->minVym};
        for (int q = 0; q < 14; q++) {
            *tabs[q] = padded(h, (size_t)Ne * N, pad);
            if (src[q]) memcpy(*tabs[q], src[q], sizeof(double) * (size_t)Ne * N);
        }
        /* drag-model tables (subgrid.c:2206-2225) and their per-cell values */
        double **dtabs[] = {&sb->allYhX, &sb->allYhY, &sb->allCn, &sb->allNmin, &sb->allNmax, &sb->allOmin,
                            &sb->allOmax};
        const double *dsrc[] = {sg->YhX, sg->YhY, sg->Cn, sg->Nmin, sg->Nmax, sg->Omin, sg->Omax};
        for (int q = 0; q < 7; q++) {
            *dtabs[q] = padded(h, (size_t)Ne * N, pad);
            if (dsrc[q]) memcpy(*dtabs[q], dsrc[q], sizeof(double) * (size_t)Ne * N);
        }
        sb->YhX = padded(h, Ng, pad);
        sb->YhY = padded(h, Ng, pad);
        sb->Cn = padded(h, Ng, pad);
        double **arr[] = {&sb->V, &sb->Z, &sb->Vxp, &sb->Vyp, &sb->Vxm, &sb->Vym, &sb->Np, &sb->Op,
            &sb->Nm, &sb->Om, &sb->Vx, &sb->Vy, &sb->Nx, &sb->Oy, &sb->Zx, &sb->Zy, &sb->Nmin,
            &sb->Omin, &sb->Nmax, &sb->Omax, &sb->minVyp, &sb->minVxp, &sb->minVym, &sb->minVxm,
            &sb->minVy, &sb->minVx};
        for (size_t q = 0; q < sizeof(arr) / sizeof(arr[0]); q++) *arr[q] = padded(h, Ng, pad);
        /* SubOutput's gather buffers (initAllSubVar allocates them, N2CI each) */
        double **outs[] = {&sb->allNx, &sb->allOy, &sb->allVx, &sb->allVy, &sb->allVo, &sb->allZo};
        for (size_t q = 0; q < sizeof(outs) / sizeof(outs[0]); q++) *outs[q] = padded(h, N, pad);
        sb->ind = ialloc(Ng);
        for (int q = 0; q < Ng; q++) sb->ind[q] = round(Ne / 2);
        computeSubArea(&h->sub, h->data, b, h->map, s);
        combineSubArea(&h->sub, h->data, b, h->map, s, 0, 1);
    }
    computeVolume(&h->data, s);
    velocityInterp(&h->data, s);
    thinLayerDrag(&h->data, s);
    findInflowLocation(&h->data, h->map, s);
    if (s->useScalar == 1) {
        for (int q = 0; q < N; q++) if (d->depth[q] <= 0) d->S[q] = 0;
        scalarMass(&h->data, s);
    }
    return h;
}

/* solveFreeSurface (nssolve.c:158-178) with the preconditioner selectable */
static void free_surface_phase(ref_handle *h)
{
    Config *s = h->setting;
    QMatrix A; Vector x, z;
    Q_Constr(&A, "A", s->N2ci, False, Rowws, Normal, True);
    V_Constr(&z, "z", s->N2ci, Normal, True);
    V_Constr(&x, "x", s->N2ci, Normal, True);
    if (h->precond == FREHD_PRECOND_SSOR) {
        solveFreeSurface(&h->data, h->sub, h->map, h->bc, h->bath, s, h->tt, A, x, z, 0, 1);
    } else {
        if (s->useSubgrid == 0) matrixCoeff(&h->data, h->map, s);
        else matrixCoeffSub(&h->data, h->map, h->sub, s);
        adjustMatrixBoundary(&h->data, h->bc, h->map, s, 0, 1);
        adjustTidalBoundary(&h->data, h->bc, h->bath, s, h->tt, 0, 1);
        setupMatrix(h->data, h->map, s, A);
        for (size_t k = 1; k <= (size_t)s->N2ci; k++) V_SetCmp(&z, k, h->data->z[k - 1]);
        V_SetAllCmp(&x, 0.0);
        SetRTCAccuracy(s->eps);
        CGIter(&A, &x, &z, s->maxIter, JacobiPrecond, 1);
        getFreeSurface(&h->data, h->map, s, x);
        zeroNegSurf(&h->data, h->bath, s, h->tt);
        enforceFreeSurfaceBC(&h->data, h->map, s);
        enforceTidalBC(&h->data, h->bath, h->bc, s, h->tt, 0, 1);
    }
    h->lastIter = GetLastNoIter();
    Q_Destr(&A); V_Destr(&x); V_Destr(&z);
}

int ref_phase(void *hv, int phase)
{
    ref_handle *h = (ref_handle *)hv;
    Config *s = h->setting;
    switch (phase) {
    case 0:
        if (h->tt + 1 > s->Nt) return FREHD_EINVAL;   /* forcing series exhausted */
        h->tt++;
        solveSourceTerm(&h->data, h->sub, h->map, h->bc, h->bath, s, h->tt, 0, 1);
        return 0;
    case 1: free_surface_phase(h); return 0;
    case 2: updateData(&h->data, &h->sub, h->map, h->bc, h->bath, s, h->tt, 0, 1); return 0;
    case 3:
        if (s->useScalar == 1)
            scalarTransport(&h->data, h->map, h->bath, h->bc, h->sub, s, h->tt, 0, 1);
        return 0;
    case 4: {
        /* monitorCFL exits the process above CFL 20; check first */
        double mx = 0;
        for (int c = 0; c < s->N2ci; c++) {
            double a = fabs(h->data->uuXP[c]) * s->dt / s->dx, b2 = fabs(h->data->vvYP[c]) * s->dt / s->dy;
            if (a > mx) mx = a;
            if (b2 > mx) mx = b2;
        }
        if (mx > 20) return FREHD_ECFL;
        monitorCFL(&h->data, h->bath, 0, s, h->tt, 0);
        return 0;
    }
    default: return FREHD_EINVAL;
    }
}

int ref_step(void *hv)
{
    if (ref_phase(hv, 0)) return FREHD_EINVAL;
    for (int p = 1; p < 4; p++) ref_phase(hv, p);
    return ref_phase(hv, 4);
}

/* The reference's callers of the step seam, around given implementations
 * of its entry points (the device library's frehdc_amd/seam/frehd_seam.c):
 * SolveAll's loop body (nssolve.c:61-104 without the output) calling
 * `step_fn` with the oneCompleteStep signature (nssolve.h:10-11) and then
 * the reference's monitorCFL; and a SolveAll replacement called as SolveAll
 * is (FREHD.c:105).  Step numbers continue from the handle's. */
typedef void (*seam_step_fn)(Data **, Sub **, Maps *, BC *, Bath *, Config *, int, QMatrix, Vector, Vector, int,
                             int);
typedef void (*seam_all_fn)(Data **, Sub **, Maps *, BC *, Bath *, Config *, int, int);

int ref_seam_steps(void *hv, void *step_fn, int nsteps)
{
    ref_handle *h = (ref_handle *)hv;
    Config *s = h->setting;
    for (int t = 0; t < nsteps; t++) {
        if (h->tt + 1 > s->Nt) return FREHD_EINVAL;
        h->tt++;
        QMatrix A; Vector z, x;
        Q_Constr(&A, "A", s->N2ci, False, Rowws, Normal, True);
        V_Constr(&z, "z", s->N2ci, Normal, True);
        V_Constr(&x, "x", s->N2ci, Normal, True);
        ((seam_step_fn)step_fn)(&h->data, &h->sub, h->map, h->bc, h->bath, s, h->tt, A, x, z, 0, 1);
        Q_Destr(&A); V_Destr(&x); V_Destr(&z);
        int rc = ref_phase(hv, 4);
        if (rc) return rc;
    }
    return 0;
}

int ref_seam_solve_all(void *hv, void *all_fn)
{
    ref_handle *h = (ref_handle *)hv;
    ((seam_all_fn)all_fn)(&h->data, &h->sub, h->map, h->bc, h->bath, h->setting, 0, 1);
    h->tt = h->setting->Nt;
    return 0;
}

int ref_run(void *hv, int nsteps)
{
    for (int t = 0; t < nsteps; t++) {
        int rc = ref_step(hv);
        if (rc) return rc;
    }
    return 0;
}

double *ref_field(void *hv, const char *name, int64_t *len)
{
    ref_handle *h = (ref_handle *)hv;
    Data *d = h->data;
    Bath *b = h->bath;
    int N = h->setting->N2ci, Ng = h->setting->N2ct;
    struct { const char *n; double *p; int64_t l; } tab[] = {
        {"surf", d->surf, Ng}, {"surfOld", d->surfOld, Ng}, {"uuXP", d->uuXP, Ng},
        {"vvYP", d->vvYP, Ng}, {"uuYP", d->uuYP, N}, {"vvXP", d->vvXP, N},
        {"depth", d->depth, N}, {"depthXP", d->depthXP, Ng}, {"depthYP", d->depthYP, Ng},
        {"EnXP", d->EnXP, Ng}, {"EnYP", d->EnYP, Ng}, {"DragXP", d->DragXP, Ng},
        {"DragYP", d->DragYP, Ng}, {"CDXP", d->CDXP, Ng}, {"CDYP", d->CDYP, Ng},
        {"Fuu", d->Fuu, Ng}, {"Fvv", d->Fvv, Ng}, {"cellV", d->cellV, N},
        {"S", d->S, Ng}, {"Sm", d->Sm, N}, {"CFLx", d->CFLx, Ng}, {"CFLy", d->CFLy, Ng},
        {"wtfXP", d->wtfXP, Ng}, {"wtfYP", d->wtfYP, Ng}, {"bottomZ", b->bottomZ, Ng},
        {"bottomXP", b->bottomXP, Ng}, {"bottomYP", b->bottomYP, Ng}, {"advX", d->advX, Ng},
        {"advY", d->advY, Ng}, {"rmvCFL", d->rmvCFL, N}, {"GnCt", d->GnCt, N},
        {"GnXP", d->GnXP, N}, {"GnXM", d->GnXM, N}, {"GnYP", d->GnYP, N}, {"GnYM", d->GnYM, N},
        {"z_transposed", d->z, N},
        {"V", h->sub->V, Ng}, {"Z", h->sub->Z, Ng}, {"Vx", h->sub->Vx, Ng}, {"Vy", h->sub->Vy, Ng},
        {"Zx", h->sub->Zx, Ng}, {"Zy", h->sub->Zy, Ng}, {"Nx", h->sub->Nx, Ng}, {"Oy", h->sub->Oy, Ng},
        {"Vxp", h->sub->Vxp, N}, {"Vxm", h->sub->Vxm, N}, {"Vyp", h->sub->Vyp, N},
        {"Vym", h->sub->Vym, N}, {"Np", h->sub->Np, N}, {"Nm", h->sub->Nm, N},
        {"Op", h->sub->Op, N}, {"Om", h->sub->Om, N}, {"minVx", h->sub->minVx, Ng},
        {"minVy", h->sub->minVy, Ng}};
    for (size_t q = 0; q < sizeof(tab) / sizeof(tab[0]); q++)
        if (strcmp(tab[q].n, name) == 0 && tab[q].p) { if (len) *len = tab[q].l; return tab[q].p; }
    if (len) *len = 0;
    return NULL;
}

int *ref_ind(void *hv) { return ((ref_handle *)hv)->sub->ind; }

int ref_tt(void *hv) { return ((ref_handle *)hv)->tt; }
int ref_pcg_iters(void *hv) { return ((ref_handle *)hv)->lastIter; }
void ref_vloss(void *hv, double out[4])
{
    Data *d = ((ref_handle *)hv)->data;
    out[0] = d->Vloss1[0]; out[1] = d->Vloss2[0]; out[2] = d->Vloss3[0]; out[3] = d->Vloss4[0];
}

void ref_destroy(void *hv)
{
    ref_handle *h = (ref_handle *)hv;
    if (!h) return;
    for (int q = 0; q < h->nblocks; q++) free(h->blocks[q]);
    /* maps, structs and small int arrays are left to process exit */
    free(h);
}

/* ---- the reference's file input/output, for the deck reader's tests ------ */
#include "fileio.h"

static Config *io_config(const char *inputFolder, const char *subgridFolder, int NX, int NY)
{
    Config *s = (Config *)calloc(1, sizeof(Config));
    snprintf(s->inputFolder, sizeof s->inputFolder, "%s", inputFolder);
    if (subgridFolder) snprintf(s->subgridFolder, sizeof s->subgridFolder, "%s", subgridFolder);
    s->NX = NX; s->NY = NY; s->nx = NX; s->ny = NY; s->np = 1;
    s->N2CI = NX * NY; s->N2ci = NX * NY; s->N2CT = (NX + 2) * (NY + 2); s->N2ct = s->N2CT;
    return s;
}

/* readOneBC (initialize.c:400-460): Nt samples of series `name` */
int ref_io_read_bc(const char *inputFolder, const char *name, int N, double tNStart, double dt, int Nt,
                   double *out)
{
    Config *s = io_config(inputFolder, NULL, 1, 1);
    s->tNStart = tNStart;
    s->dt = dt;
    s->Nt = Nt;
    s->tNEnd = tNStart + dt * Nt;
    readOneBC(out, (char *)name, s, N);
    free(s);
    return 0;
}

/* ReadBathymetry + InitBathymetry (bathymetry.c:19-104) */
int ref_io_bathymetry(const char *inputFolder, const char *subgridFolder, int useSubgrid, int NX, int NY,
                      double *allZ, double *allEdgeX, double *allEdgeY, double *offset)
{
    Config *s = io_config(inputFolder, subgridFolder, NX, NY);
    s->useSubgrid = useSubgrid;
    Bath *b = NULL;
    ReadBathymetry(&b, s, 0);
    InitBathymetry(b, s, 0);
    memcpy(allZ, b->allZ, sizeof(double) * NX * NY);
    if (useSubgrid) {
        memcpy(allEdgeX, b->allEdgeX, sizeof(double) * NX * NY);
        memcpy(allEdgeY, b->allEdgeY, sizeof(double) * NX * NY);
    }
    *offset = b->offset[0];
    free(s);
    return 0;
}

/* initAllSubVar (subgrid.c:2112-2284): one level-major table (0..13 = V Z
 * Vxp Vxm Vyp Vym Np Nm Op Om minVxp minVxm minVyp minVym) */
int ref_io_subtables(const char *inputFolder, const char *subgridFolder, int NX, int NY, double surfmin,
                     double surfmax, double dsurf, double offset, int useminV, double *out[14], int *Ne)
{
    Config *s = io_config(inputFolder, subgridFolder, NX, NY);
    s->useSubgrid = 1;
    s->surfmin = surfmin;
    s->surfmax = surfmax;
    s->dsurf = dsurf;
    s->useminV = useminV;
    s->useminA = 0;
    s->useSubDrag = 0;
    Bath b;
    memset(&b, 0, sizeof b);
    double off = offset;
    b.offset = &off;
    Sub *sub = NULL;
    initAllSubVar(&sub, &b, s, 0);
    *Ne = sub->Ne;
    double *src[14] = {sub->allV, sub->allZ, sub->allVxp, sub->allVxm, sub->allVyp, sub->allVym, sub->allNp,
                       sub->allNm, sub->allOp, sub->allOm, sub->allminVxp, sub->allminVxm, sub->allminVyp,
                       sub->allminVym};
    for (int t = 0; t < 14; t++)
        if (out[t] && src[t]) memcpy(out[t], src[t], sizeof(double) * NX * NY * sub->Ne);
    free(s);
    return 0;
}

/* initAllSubVar's optional tables, level-major: YhX, YhY with useSubDrag =
 * 2, Cn with 3 or 4 (subgrid.c:2206-2225); Nmin, Nmax, Omin, Omax with
 * useminA = 1 (:2160-2190) */
int ref_io_extratables(const char *inputFolder, const char *subgridFolder, int NX, int NY, double surfmin,
                       double surfmax, double dsurf, double offset, int useSubDrag, int useminA, double *out[7],
                       int *Ne)
{
    Config *s = io_config(inputFolder, subgridFolder, NX, NY);
    s->useSubgrid = 1;
    s->surfmin = surfmin;
    s->surfmax = surfmax;
    s->dsurf = dsurf;
    s->useminV = 0;
    s->useminA = useminA;
    s->useSubDrag = useSubDrag;
    Bath b;
    memset(&b, 0, sizeof b);
    double off = offset;
    b.offset = &off;
    Sub *sub = NULL;
    initAllSubVar(&sub, &b, s, 0);
    *Ne = sub->Ne;
    double *src[7] = {useSubDrag == 2 ? sub->allYhX : NULL, useSubDrag == 2 ? sub->allYhY : NULL,
                      useSubDrag >= 3 ? sub->allCn : NULL, useminA ? sub->allNmin : NULL,
                      useminA ? sub->allNmax : NULL, useminA ? sub->allOmin : NULL, useminA ? sub->allOmax : NULL};
    for (int t = 0; t < 7; t++)
        if (out[t] && src[t]) memcpy(out[t], src[t], sizeof(double) * NX * NY * sub->Ne);
    free(s);
    return 0;
}

/* WriteFile (fileio.c:181-206) */
int ref_io_write(const char *saveFolder, const char *name, int tt, const double *arr, int N)
{
    Config *s = io_config("", NULL, 1, 1);
    snprintf(s->saveFolder, sizeof s->saveFolder, "%s", saveFolder);
    WriteFile((double *)arr, (char *)name, s, tt, N);
    free(s);
    return 0;
}
