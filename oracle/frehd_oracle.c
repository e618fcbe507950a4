/*
 * frehd_oracle.c -- CPU restatement of one FrehdC v4.2 Casulli time step.
 *
 * TEST INFRASTRUCTURE ONLY (see frehd_oracle.h): the checker for the device
 * library, never part of it.  Build: oracle/Makefile (gcc -O2
 * -ffp-contract=off, so no fused multiply-add changes a rounding).
 *
 * Scope: the plain-grid and subgrid paths with first-order upwind or TVD
 * superbee scalar transport, Manning or constant drag, optional thin-layer
 * drag and wind stress, no evaporation/rain (their flags are never set by
 * the reference's reader), one strip (np = 1, irank = 0 = nrank - 1).
 *
 * Order of floating-point operations follows the reference expression by
 * expression (C evaluates a*b*c as (a*b)*c and a+b+c as (a+b)+c), so the
 * result matches the reference bit for bit; tests/test_oracle_vs_reference.py
 * checks that against the reference's own functions compiled into
 * oracle/_ref.  Where the reference reads outside an array, the value read
 * is defined here as 0.0 and the reference harness
 * pads every array with zeros so both agree:
 *   - velocityInterp reads uuXP[ii-1] at ii = 0 and vvYP[ii-nx] for ii < nx
 *     (nsfunctions.c:342-345);
 *   - limiterCFL reads depth[] at ghost indices although depth holds N values
 *     (nsfunctions.c:462-465); depth at a ghost index is 0 here.
 * CFLx/CFLy are read by the first step's advectionTerm before monitorCFL ever
 * wrote them (uninitialised malloc in initialize.c:222-223): 0 here.
 *
 * Parity pins: every function below is pinned by the reference harness
 * comparison (bit-exact per phase and over multi-step runs).  The CG is
 * additionally pinned by a dense solve of the same 5-point system
 * (tests/test_oracle_math.py); the Jacobi variant is pinned the same way.
 */
#include "frehd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* LASpack's tolerance tests (laspack/elcmp.h:31-32). */
#define LAS_IS_ZERO(a) (fabs(a) < 10.0 * 2.2250738585072014e-308)
#define LAS_IS_ONE(a)  (fabs((a) - 1.0) < 10.0 * 2.220446049250313e-16)

struct fo_state {
    frehd_config cfg;
    int nx, ny, N, Ng;
    int tt;
    /* forcing (owned copies) */
    int32_t *tideLocP, *tideLocM, *inflowLoc;
    double *tideP, *tideM, *inflow, *tidalPS, *tidalMS, *manningCell, *windspd, *winddir;
    /* Ng arrays */
    double *surf, *surfOld, *uuXP, *vvYP, *Fuu, *Fvv, *depthXP, *depthYP;
    double *EnXP, *EnYP, *DragXP, *DragYP, *CDXP, *CDYP, *wtfXP, *wtfYP;
    double *S, *CFLx, *CFLy, *advX, *advY, *bottomZ, *bottomXP, *bottomYP;
    /* N arrays (allocated Ng so ghost-index reads are defined zeros) */
    double *uuYP, *vvXP, *depth, *cellV, *Sm, *rmvCFL;
    double *GnCt, *GnXP, *GnXM, *GnYP, *GnYM, *z;
    double Vloss[4];
    int pcgIters;
    long long pcgItersTotal;
    double relres;
    /* subgrid (Sub, subgrid.h:11-37): tables level-major [k*N + ii] */
    int sub, Ne;
    double *allSurf, *edgeX, *edgeY;
    double *tV, *tZ, *tVxp, *tVxm, *tVyp, *tVym, *tNp, *tNm, *tOp, *tOm;
    double *tminVxp, *tminVxm, *tminVyp, *tminVym;
    double *tYhX, *tYhY, *tCn, *YhX, *YhY, *Cn;   /* useSubDrag 2 / 3, 4 */
    double *tNmin, *tNmax, *tOmin, *tOmax;          /* useminA = 1 */
    double *V, *Z, *Vxp, *Vyp, *Vxm, *Vym, *Np, *Nm, *Op, *Om, *Vx, *Vy, *Zx, *Zy, *Nx, *Oy;
    double *minVxp, *minVxm, *minVyp, *minVym, *minVx, *minVy, *Nmin, *Omin, *Nmax, *Omax;
    int *ind;
};

/* ---- grid index helpers (map.c:15-128) ------------------------------------ */
static inline int ci(const fo_state *s, int c) { return c % s->nx; }
static inline int cj(const fo_state *s, int c) { return c / s->nx; }
static inline int gjP(const fo_state *s, int i) { return s->N + i; }
static inline int gjM(const fo_state *s, int i) { return s->N + s->nx + i; }
static inline int giP(const fo_state *s, int j) { return s->N + 2 * s->nx + j; }
static inline int giM(const fo_state *s, int j) { return s->N + 2 * s->nx + s->ny + j; }
static inline int nbP(const fo_state *s, int c) { int i = ci(s, c); return i == s->nx - 1 ? giP(s, cj(s, c)) : c + 1; }
static inline int nbM(const fo_state *s, int c) { int i = ci(s, c); return i == 0 ? giM(s, cj(s, c)) : c - 1; }
static inline int nbN(const fo_state *s, int c) { int j = cj(s, c); return j == s->ny - 1 ? gjP(s, ci(s, c)) : c + s->nx; }
static inline int nbS(const fo_state *s, int c) { int j = cj(s, c); return j == 0 ? gjM(s, ci(s, c)) : c - s->nx; }
/* two cells away (createMaps' iPPjc, iMMjc, icjPP, icjMM, map.c:68-98): the
   ghost beyond the edge where the second cell falls outside; icjMM of row 1,
   which map.c:69-74 never sets, is the jM ghost likewise */
static inline int nbPP(const fo_state *s, int c) { int i = ci(s, c); return i >= s->nx - 2 ? giP(s, cj(s, c)) : c + 2; }
static inline int nbMM(const fo_state *s, int c) { int i = ci(s, c); return i <= 1 ? giM(s, cj(s, c)) : c - 2; }
static inline int nbNN(const fo_state *s, int c) { int j = cj(s, c); return j >= s->ny - 2 ? gjP(s, ci(s, c)) : c + 2 * s->nx; }
static inline int nbSS(const fo_state *s, int c) { int j = cj(s, c); return j <= 1 ? gjM(s, ci(s, c)) : c - 2 * s->nx; }
/* corners, map.c:101-104 */
static inline int cornerIMJP(const fo_state *s) { return s->Ng - 4; }
static inline int cornerIPJP(const fo_state *s) { return s->Ng - 3; }
static inline int cornerIPJM(const fo_state *s) { return s->Ng - 2; }
static inline int cornerIMJM(const fo_state *s) { return s->Ng - 1; }
/* flat read with a zero for indices below the array (velocityInterp quirk) */
static inline double rd0(const double *a, long idx) { return idx < 0 ? 0.0 : a[idx]; }

static double *zalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }

/* ---- boundary fills ------------------------------------------------------- */
/* enforceFreeSurfaceBC, nsfunctions.c:317-334 */
static void fill_surface_ghosts(fo_state *s, double *a)
{
    const int nx = s->nx, ny = s->ny, N = s->N;
    for (int i = 0; i < nx; i++) { a[gjP(s, i)] = a[N - nx + i]; a[gjM(s, i)] = a[i]; }
    for (int j = 0; j < ny; j++) { a[giP(s, j)] = a[j * nx + nx - 1]; a[giM(s, j)] = a[j * nx]; }
    a[cornerIPJP(s)] = a[giP(s, ny - 1)];
    a[cornerIPJM(s)] = a[giM(s, ny - 1)];
    a[cornerIMJP(s)] = a[giP(s, 0)];
    a[cornerIMJM(s)] = a[giM(s, 0)];
}

/* enforceVelocityBC, nsfunctions.c:289-314 */
static void fill_velocity_ghosts(fo_state *s)
{
    const int nx = s->nx, ny = s->ny, N = s->N;
    double *u = s->uuXP, *v = s->vvYP;
    for (int i = 0; i < nx; i++) {
        u[gjP(s, i)] = u[N - nx + i]; u[gjM(s, i)] = u[i];
        v[gjP(s, i)] = v[N - nx + i]; v[gjM(s, i)] = v[i];
    }
    for (int j = 0; j < ny; j++) {
        u[giP(s, j)] = u[j * nx + nx - 1]; u[giM(s, j)] = u[j * nx];
        v[giP(s, j)] = v[j * nx + nx - 1]; v[giM(s, j)] = v[j * nx];
    }
    u[cornerIPJP(s)] = u[gjP(s, nx - 1)];
    u[cornerIPJM(s)] = u[gjM(s, nx - 1)];
    u[cornerIMJP(s)] = u[gjP(s, 0)];
    u[cornerIMJM(s)] = u[gjM(s, 0)];
    v[cornerIPJP(s)] = v[giP(s, ny - 1)];
    v[cornerIPJM(s)] = v[giM(s, ny - 1)];
    v[cornerIMJP(s)] = v[giP(s, 0)];
    v[cornerIMJM(s)] = v[giM(s, 0)];
}

/* ---- Init pieces ----------------------------------------------------------- */
/* enforceBathBC (initialize.c:158-175) + updateAllFaceElev for useCellEdge = 0,
   useSubgrid = 0 (nsfunctions.c:76-140) */
static void init_bathymetry(fo_state *s)
{
    const int nx = s->nx, ny = s->ny, N = s->N;
    double *zb = s->bottomZ, *bx = s->bottomXP, *by = s->bottomYP;
    for (int i = 0; i < nx; i++) { zb[gjM(s, i)] = zb[i]; zb[gjP(s, i)] = zb[N - nx + i]; }
    for (int j = 0; j < ny; j++) { zb[giM(s, j)] = zb[j * nx]; zb[giP(s, j)] = zb[j * nx + nx - 1]; }
    for (int c = 0; c < s->Ng; c++) { bx[c] = 0.0; by[c] = 0.0; }
    if (s->sub) { for (int c = 0; c < N; c++) bx[c] = s->edgeX[c]; }   /* useSubgrid != 0 */
    else for (int c = 0; c < N; c++) { double a = zb[c], b = zb[nbP(s, c)]; bx[c] = a > b ? a : b; }
    for (int j = 0; j < ny; j++) {
        bx[giM(s, j)] = zb[j * nx];
        bx[giP(s, j)] = bx[j * nx + nx - 1];
    }
    for (int i = 0; i < nx; i++) { bx[gjM(s, i)] = bx[i]; bx[gjP(s, i)] = bx[N - nx + i]; }
    if (s->sub) { for (int c = 0; c < N; c++) by[c] = s->edgeY[c]; }
    else for (int c = 0; c < N; c++) { double a = zb[c], b = zb[nbN(s, c)]; by[c] = a > b ? a : b; }
    for (int i = 0; i < nx; i++) {
        double a = zb[gjM(s, i)], b = zb[i];
        by[gjM(s, i)] = a > b ? a : b;
        by[gjP(s, i)] = by[N - nx + i];
    }
    for (int j = 0; j < ny; j++) { by[giM(s, j)] = by[j * nx]; by[giP(s, j)] = by[j * nx + nx - 1]; }
}

/* ---- shared pieces of Init and updateData ---------------------------------- */
/* zeroNegSurf, nsfunctions.c:436-449 */
static void zero_neg_surf(fo_state *s, int tt)
{
    for (int c = 0; c < s->N; c++) {
        if (s->surf[c] < s->bottomZ[c]) {
            double diff = s->bottomZ[c] - s->surf[c];
            s->surf[c] = s->bottomZ[c];
            if (tt > 1) s->Vloss[2] -= diff * s->cfg.dx * s->cfg.dy;
        }
    }
}

/* enforceTidalBC, nsfunctions.c:412-433 */
static void tidal_surface(fo_state *s, int tt)
{
    const double off = s->cfg.offset;
    if (s->cfg.bcType != 2)
        for (int k = 0; k < s->cfg.tideLocLengthP; k++) {
            int c = s->tideLocP[k];
            if (s->tideP[tt] + off > s->bottomZ[c]) s->surf[c] = s->tideP[tt] + off;
        }
    if (s->cfg.bcType != 1)
        for (int k = 0; k < s->cfg.tideLocLengthM; k++) {
            int c = s->tideLocM[k];
            if (s->tideM[tt] + off > s->bottomZ[c]) s->surf[c] = s->tideM[tt] + off;
        }
}

/* updateDepth, nsfunctions.c:239-247 */
static void update_depth(fo_state *s)
{
    for (int c = 0; c < s->N; c++) {
        double d = s->surf[c] - s->bottomZ[c];
        s->depth[c] = d < 0 ? 0 : d;
    }
}

/* updateAllFaceDepth, nsfunctions.c:143-236 (irank = 0 = nrank - 1) */
static void update_face_depth(fo_state *s)
{
    const int nx = s->nx, ny = s->ny, N = s->N;
    double *h = s->surf, *dX = s->depthXP, *dY = s->depthYP;
    const double *bX = s->bottomXP, *bY = s->bottomYP;
    for (int c = 0; c < s->Ng; c++) { dX[c] = 0; dY[c] = 0; }
    for (int c = 0; c < N; c++) {
        double a = h[c], b = h[nbP(s, c)];
        dX[c] = (a > b ? a : b) - bX[c];
    }
    for (int j = 0; j < ny; j++) {
        dX[giM(s, j)] = h[j * nx] - bX[giM(s, j)];
        dX[giP(s, j)] = h[giP(s, j)] - bX[giP(s, j)];
        dX[j * nx + nx - 1] = 0;
        dX[giP(s, j)] = 0;
        dX[giM(s, j)] = 0;
    }
    for (int i = 0; i < nx; i++) {
        int gm = gjM(s, i), gp = gjP(s, i);
        dX[gm] = (h[gm] > h[gm + 1] ? h[gm] : h[gm + 1]) - bX[gm];
        dX[gp] = (h[gp] > h[gp + 1] ? h[gp] : h[gp + 1]) - bX[gp];
    }
    for (int c = 0; c < N; c++) {
        double a = h[c], b = h[nbN(s, c)];
        dY[c] = (a > b ? a : b) - bY[c];
    }
    for (int i = 0; i < nx; i++) {
        int gm = gjM(s, i), gp = gjP(s, i);
        dY[gm] = (h[gm] > h[i] ? h[gm] : h[i]) - bY[gm];
        dY[gp] = h[gp] - bY[gp];
        dY[gm] = 0;
    }
    for (int j = 0; j < ny; j++) { dY[giM(s, j)] = dY[j * nx]; dY[giP(s, j)] = dY[j * nx + nx - 1]; }
    for (int i = 0; i < nx; i++) { dY[N - nx + i] = 0; dY[gjP(s, i)] = 0; }
    for (int i = 0; i < nx; i++) dY[gjM(s, i)] = 0;
    for (int c = 0; c < s->Ng; c++) { if (dX[c] < 0) dX[c] = 0; if (dY[c] < 0) dY[c] = 0; }
}

/* computeVolume, nsfunctions.c:361-370 */
static void compute_volume(fo_state *s)
{
    for (int c = 0; c < s->N; c++) {
        s->cellV[c] = s->cfg.dx * s->cfg.dy * s->depth[c];
        if (s->cfg.useScalar == 1)
            for (int q = 0; q < s->cfg.nScalar; q++)
                s->Sm[q * s->Ng + c] = s->S[q * s->Ng + c] * s->cellV[c];
    }
}

/* velocityInterp, nsfunctions.c:337-347 (flat-index neighbours) */
static void velocity_interp(fo_state *s)
{
    const long nx = s->nx;
    for (long c = 0; c < s->N; c++) {
        s->uuYP[c] = 0.25 * (s->uuXP[c] + rd0(s->uuXP, c - 1) + s->uuXP[c + nx] + rd0(s->uuXP, c - 1 + nx));
        s->vvXP[c] = 0.25 * (s->vvYP[c] + s->vvYP[c + 1] + rd0(s->vvYP, c - nx) + rd0(s->vvYP, c + 1 - nx));
    }
}

static inline double manning_coeff(const fo_state *s, int c)
{
    double n = s->manningCell ? s->manningCell[c] : s->cfg.manningN;
    return s->cfg.g * n * n;
}

/* thinLayerDrag, nsfunctions.c:691-727 */
static void thin_layer_drag(fo_state *s)
{
    const frehd_config *k = &s->cfg;
    if (k->CDnotN == 1) {
        for (int c = 0; c < s->Ng; c++) {
            if (s->depthXP[c] <= k->hD)
                s->CDXP[c] = k->CDx + (k->CDmax - k->CDx) * (k->hD - s->depthXP[c]) / k->hD;
            if (s->depthYP[c] <= k->hD)
                s->CDYP[c] = k->CDy + (k->CDmax - k->CDy) * (k->hD - s->depthYP[c]) / k->hD;
        }
    } else {
        for (int c = 0; c < s->N; c++) {
            if (s->depth[c] > 0) {
                double coeff = manning_coeff(s, c);
                if (s->depthXP[c] <= k->hD) s->CDXP[c] = coeff / pow(s->depth[c], 2.0 / 3.0);
                if (s->depthYP[c] <= k->hD) s->CDYP[c] = coeff / pow(s->depth[c], 2.0 / 3.0);
            }
        }
    }
}

/* ---- subgrid runtime (subgrid.c:55-526) ----------------------------------- */
/* interpSurf, utilities.c:155-162 */
static inline double interp_surf(double x, double x1, double y1, double x2, double y2)
{
    if (x2 == x1) return y1;
    return (y2 - y1) * (x - x2) / (x2 - x1) + y2;
}

/* computeSubArea, subgrid.c:55-251: table lookup per cell.  edgeX/edgeY hold
   N values; the west/south neighbour of a boundary cell is a ghost index,
   read as 0 here (the reference reads past the array). */
static void compute_sub_area(fo_state *s)
{
    const frehd_config *k = &s->cfg;
    const int N = s->N, Ne = s->Ne;
    const double dsurf = k->dsurf;
    for (int c = 0; c < N; c++) {
        const double h = s->surf[c];
        while (fabs(s->allSurf[s->ind[c]] - h) > 0.5 * dsurf + dsurf / 100) {
            if (h > s->allSurf[s->ind[c]]) s->ind[c] += 1;
            else s->ind[c] -= 1;
            if (s->ind[c] < 0) { s->ind[c] = 0; break; }
            if (s->ind[c] > Ne - 1) { s->ind[c] = Ne - 1; break; }
        }
        const int j1 = s->ind[c];
        int j2;
        if (j1 == 0) j2 = 0;
        else if (j1 == Ne - 1) j2 = Ne - 1;
        else j2 = h > s->allSurf[j1] ? j1 + 1 : j1 - 1;
        const size_t k1 = (size_t)N * j1 + c, k2 = (size_t)N * j2 + c;
        const double s1 = s->allSurf[j1], s2 = s->allSurf[j2];
        /* drag-model values at the lower level, subgrid.c:121-127 */
        if (k->useSubDrag == 2) { s->YhX[c] = s->tYhX[k1]; s->YhY[c] = s->tYhY[k1]; }
        if (k->useSubDrag == 4 || k->useSubDrag == 3) s->Cn[c] = s->tCn[k1];
        s->V[c] = interp_surf(h, s1, s->tV[k1], s2, s->tV[k2]);
        s->Z[c] = interp_surf(h, s1, s->tZ[k1], s2, s->tZ[k2]);
        s->Vxp[c] = interp_surf(h, s1, s->tVxp[k1], s2, s->tVxp[k2]);
        s->Vyp[c] = interp_surf(h, s1, s->tVyp[k1], s2, s->tVyp[k2]);
        s->Vxm[c] = interp_surf(h, s1, s->tVxm[k1], s2, s->tVxm[k2]);
        s->Vym[c] = interp_surf(h, s1, s->tVym[k1], s2, s->tVym[k2]);
        s->minVxp[c] = interp_surf(h, s1, s->tminVxp[k1], s2, s->tminVxp[k2]);
        s->minVyp[c] = interp_surf(h, s1, s->tminVyp[k1], s2, s->tminVyp[k2]);
        s->minVxm[c] = interp_surf(h, s1, s->tminVxm[k1], s2, s->tminVxm[k2]);
        s->minVym[c] = interp_surf(h, s1, s->tminVym[k1], s2, s->tminVym[k2]);
        /* Nmin and Omin, subgrid.c:210-220 */
        if (k->useminA != 0) {
            s->Nmin[c] = interp_surf(h, s1, s->tNmin[k1], s2, s->tNmin[k2]);
            s->Omin[c] = interp_surf(h, s1, s->tOmin[k1], s2, s->tOmin[k2]);
            s->Nmax[c] = interp_surf(h, s1, s->tNmax[k1], s2, s->tNmax[k2]);
            s->Omax[c] = interp_surf(h, s1, s->tOmax[k1], s2, s->tOmax[k2]);
        }
        const int w = nbM(s, c), so = nbS(s, c);
        const double wd[4] = {s->edgeX[c], w < N ? s->edgeX[w] : 0.0, s->edgeY[c], so < N ? s->edgeY[so] : 0.0};
        const double *tab[4] = {s->tNp, s->tNm, s->tOp, s->tOm};
        double *out[4] = {s->Np, s->Nm, s->Op, s->Om};
        for (int f = 0; f < 4; f++) {
            if (wd[f] > s1 && wd[f] < s2) {
                if (wd[f] > h) out[f][c] = 0;
                else out[f][c] = interp_surf(h, wd[f], 0.0, s2, tab[f][k2]);
            } else {
                out[f][c] = interp_surf(h, s1, tab[f][k1], s2, tab[f][k2]);
            }
        }
        if (s->depth[c] == 0) { s->Vxp[c] = 0; s->Vyp[c] = 0; s->Vxm[c] = 0; s->Vym[c] = 0; }
        if (h - s->bottomXP[c] <= 0) s->Np[c] = 0;
        if (h - s->bottomYP[c] <= 0) s->Op[c] = 0;
        if (w < N && h - s->bottomXP[w] <= 0) s->Nm[c] = 0;
        if (so < N && h - s->bottomYP[so] <= 0) s->Om[c] = 0;
    }
}

/* the effective-volume factor of combineSubArea, subgrid.c:377-403 */
static double volume_lambda(double r, double cap)
{
    double lambda;
    if (r > 0 && r < 1) {
        lambda = -5.01 * pow(1 - r, 2) + 6.35 * (1 - r) - 0.6;
        if (1 - r > 0.67) lambda = cap;
        else if (lambda < 0) lambda = 0.0;
    } else {
        lambda = 0.0;
    }
    return lambda;
}

/* combineSubArea, subgrid.c:254-526 (useminV = 1's adjustment needs
   Nmin > 0, i.e. useminA = 1) */
static void combine_sub_area(fo_state *s)
{
    const frehd_config *k = &s->cfg;
    const int N = s->N, nx = s->nx, ny = s->ny;
    const double dxdy = k->dx * k->dy;
    for (int c = 0; c < N; c++) {
        const int e = nbP(s, c), n = nbN(s, c);
        if (e < N) s->Nx[c] = s->Np[c] >= s->Nm[e] ? s->Np[c] : s->Nm[e];
        else s->Nx[c] = s->Np[c];
        if (n < N) s->Oy[c] = s->Op[c] >= s->Om[n] ? s->Op[c] : s->Om[n];
        else s->Oy[c] = s->Op[c];
        if (e < N) {
            if (s->Vxp[c] > 0 && s->Vxm[e] > 0) s->Vx[c] = s->Vxp[c] + s->Vxm[e];
            else s->Vx[c] = 2.0 * (s->Vxp[c] + s->Vxm[e]);
            s->Zx[c] = s->Z[c];
            if (s->Vx[c] > 0 && s->Zx[c] <= 0) s->Zx[c] = dxdy;
            if (k->useminV != 0) s->minVx[c] = s->minVxp[c] + s->minVxm[e];
            if (s->surf[c] < s->bottomXP[c]) { s->Vx[c] = s->Vxm[e]; s->Zx[c] = 0.5 * s->Z[c]; }
            if (s->surf[e] < s->bottomXP[c]) { s->Vx[c] = s->Vxp[c]; s->Zx[c] = 0.5 * s->Z[c]; }
        } else {
            s->Vx[c] = 2 * s->Vxp[c];
            s->Zx[c] = 2 * s->Z[c];
            s->minVx[c] = 2 * s->minVxp[c];
        }
        if (n < N) {
            if (s->Vyp[c] > 0 && s->Vym[n] > 0) s->Vy[c] = s->Vyp[c] + s->Vym[n];
            else s->Vy[c] = 2.0 * (s->Vyp[c] + s->Vym[n]);
            s->Zy[c] = s->Z[c];
            if (s->Vy[c] > 0 && s->Zy[c] <= 0) s->Zy[c] = dxdy;
            if (k->useminV != 0) s->minVy[c] = s->minVyp[c] + s->minVym[n];
            if (s->surf[c] < s->bottomYP[c]) { s->Vy[c] = s->Vym[n]; s->Zy[c] = 0.5 * s->Z[c]; }
            if (s->surf[n] < s->bottomYP[c]) { s->Vy[c] = s->Vyp[c]; s->Zy[c] = 0.5 * s->Z[c]; }
        } else {
            s->Vy[c] = 2 * s->Vyp[c];
            s->Zy[c] = 2 * s->Z[c];
            s->minVy[c] = 2 * s->minVyp[c];
        }
        /* useminA (the reference's local beta = 0 selects the first branch),
           subgrid.c:360-373 */
        const double epsi = 0.0001;
        if (k->useminA == 1) {
            if ((s->Nx[c] - s->Nmax[c] > epsi) && (s->Nmax[c] > 0))
                s->Nx[c] = s->Nx[c] - k->beta * (s->Nx[c] - s->Nmax[c]);
            if ((s->Oy[c] - s->Omax[c] > epsi) && (s->Omax[c] > 0))
                s->Oy[c] = s->Oy[c] - k->beta * (s->Oy[c] - s->Omax[c]);
        }
        /* useminV, subgrid.c:375-414 */
        if (k->useminV == 1) {
            const double Vfloor = 1.0;
            if ((s->Nmin[c] > 0) && (s->Vx[c] > s->minVx[c]) && (s->minVx[c] > 0)) {
                double r = s->Nmin[c] / s->Nmax[c];
                double lambda = volume_lambda(r, 1.35);
                s->Vx[c] = s->Vx[c] - (1 - r) * lambda * s->Vx[c];
                if (s->Vx[c] < Vfloor) s->Vx[c] = Vfloor;
            }
            if ((s->Omin[c] > 0) && (s->Vy[c] > s->minVy[c]) && (s->minVy[c] > 0)) {
                double r = s->Omin[c] / s->Omax[c];
                double lambda = volume_lambda(r, 1.4);
                s->Vy[c] = s->Vy[c] - (1 - r) * lambda * s->Vy[c];
                if (s->Vy[c] < Vfloor) s->Vy[c] = Vfloor;
            }
        } else if (k->useminV == 2) {
            s->Vx[c] = s->minVx[c];
            s->Vy[c] = s->minVy[c];
        }
    }
    const double Vmin = 0.002, Amin = 0.002, Azmin = 0.002;
    for (int c = 0; c < N; c++) {
        if (s->Vx[c] < Vmin) s->Vx[c] = 0.0;
        if (s->Vy[c] < Vmin) s->Vy[c] = 0.0;
        if (s->V[c] < Vmin) s->V[c] = 0.0;
        if (s->Z[c] < Azmin) s->Z[c] = 1.0;
        if (s->Zx[c] < Amin) s->Zx[c] = 0.0;
        if (s->Zy[c] < Amin) s->Zy[c] = 0.0;
        if (s->Nx[c] < Amin) s->Nx[c] = 0.0;
        if (s->Oy[c] < Amin) s->Oy[c] = 0.0;
    }
    for (int j = 0; j < ny; j++) {
        const int gm = giM(s, j), gp = giP(s, j), bm = j * nx, bp = j * nx + nx - 1;
        s->Vx[gm] = s->Vx[bm]; s->Vy[gm] = s->Vy[bm]; s->Zx[gm] = s->Zx[bm]; s->Zy[gm] = s->Zy[bm];
        s->Nx[gm] = s->Nx[bm]; s->Oy[gm] = s->Oy[bm]; s->V[gm] = s->V[bm];
        s->minVx[gm] = s->minVy[bm]; s->minVy[gm] = s->minVx[bm]; s->Z[gm] = s->Z[bm];
        s->Vx[gp] = s->Vx[bp]; s->Vy[gp] = s->Vy[bp]; s->Zx[gp] = s->Zx[bp]; s->Zy[gp] = s->Zy[bp];
        s->Nx[gp] = s->Nx[bp]; s->Oy[gp] = s->Oy[bp]; s->V[gp] = s->V[bp];
        s->minVy[gp] = s->minVx[bp]; s->minVx[gp] = s->minVy[bp]; s->Z[gp] = s->Z[bp];
        s->Nx[gm] = 0; s->Nx[gp] = 0; s->Nx[bp] = 0;
    }
    for (int i = 0; i < nx; i++) {
        const int gm = gjM(s, i), gp = gjP(s, i), bm = i, bp = N - nx + i;
        s->Vx[gm] = s->Vx[bm]; s->Vy[gm] = s->Vy[bm]; s->Zx[gm] = s->Zx[bm]; s->Zy[gm] = s->Zy[bm];
        s->Nx[gm] = s->Nx[bm]; s->Oy[gm] = s->Oy[bm]; s->V[gm] = s->V[bm];
        s->minVx[gm] = s->minVx[bm]; s->minVy[gm] = s->minVy[bm]; s->Z[gm] = s->Z[bm];
        s->Vx[gp] = s->Vx[bp]; s->Vy[gp] = s->Vy[bp]; s->Zx[gp] = s->Zx[bp]; s->Zy[gp] = s->Zy[bp];
        s->Nx[gp] = s->Nx[bp]; s->Oy[gp] = s->Oy[bp]; s->V[gp] = s->V[bp];
        s->minVx[gp] = s->minVx[bp]; s->minVy[gp] = s->minVy[bp]; s->Z[gp] = s->Z[bp];
        s->Oy[gm] = 0;
    }
    /* YP faces closed except at tidal cells (subgrid.c:508-524) */
    if (k->bcType != 2) {
        for (int i = 0; i < nx; i++) {
            int tidal = 0;
            for (int q = 0; q < k->tideLocLengthP; q++)
                if (s->tideLocP[q] - nx * (ny - 1) == i) tidal = 1;
            if (!tidal) { s->Oy[N - nx + i] = 0; s->Oy[gjP(s, i)] = 0; }
        }
    }
}

/* ---- Phase P: solveSourceTerm (nssolve.c:121-155), plain grid -------------- */
static void phase_source(fo_state *s)
{
    const frehd_config *k = &s->cfg;
    const int N = s->N;
    const double dx = k->dx, dy = k->dy, dt = k->dt;
    double *u = s->uuXP, *v = s->vvYP, *uY = s->uuYP, *vX = s->vvXP;
    double *aX = s->advX, *aY = s->advY;

    /* explicitVelocity, nsfunctions.c:498-507 */
    for (int c = 0; c < N; c++) { s->EnXP[c] = u[c]; s->EnYP[c] = v[c]; }

    /* advectionTerm, nsfunctions.c:510-586: first-order upwind */
    for (int c = 0; c < N; c++) {
        int e = nbP(s, c), w = nbM(s, c), n = nbN(s, c), so = nbS(s, c);
        double uc = u[c], vx = vX[c];
        aX[c] = (0.5 / dx) * (uc + fabs(uc)) * (uc - u[w])
              + (0.5 / dx) * (uc - fabs(uc)) * (u[e] - uc)
              + (0.5 / dy) * (vx + fabs(vx)) * (uc - u[so])
              + (0.5 / dy) * (vx - fabs(vx)) * (u[n] - uc);
        double vc = v[c], uy = uY[c];
        aY[c] = (0.5 / dx) * (uy + fabs(uy)) * (vc - v[w])
              + (0.5 / dx) * (uy - fabs(uy)) * (v[e] - vc)
              + (0.5 / dy) * (vc + fabs(vc)) * (vc - v[so])
              + (0.5 / dy) * (vc - fabs(vc)) * (v[n] - vc);
        if (uc == 0) aX[c] = 0;
        if (vc == 0) aY[c] = 0;
    }
    /* CFL limiter applied in place, cell by cell in increasing index: a cell
       rescales its own term and its four neighbours' (nsfunctions.c:536-585);
       the y branch's zeroing tests CFLx (nsfunctions.c:574) */
    const double lo = k->CFLl, hi = k->CFLh;
    for (int c = 0; c < N; c++) {
        int t[5] = {c, nbP(s, c), nbM(s, c), nbN(s, c), nbS(s, c)};
        double cx = s->CFLx[c], cy = s->CFLy[c];
        if (cx > lo && cx <= hi) {
            for (int q = 0; q < 5; q++) aX[t[q]] = aX[t[q]] * (hi - s->CFLx[t[q]]) / (hi - lo);
        } else if (cx > hi) {
            for (int q = 0; q < 5; q++) aX[t[q]] = 0;
        }
        if (cy > lo && cy < hi) {
            for (int q = 0; q < 5; q++) aY[t[q]] = aY[t[q]] * (hi - s->CFLy[t[q]]) / (hi - lo);
        } else if (cx > hi) {
            for (int q = 0; q < 5; q++) aY[t[q]] = 0;
        }
        s->EnXP[c] = s->EnXP[c] - dt * aX[c];
        s->EnYP[c] = s->EnYP[c] - dt * aY[c];
    }

    if (s->sub) {
        /* diffusionTermSub, subgrid.c:781-815 */
        for (int c = 0; c < N; c++) {
            int e = nbP(s, c), w = nbM(s, c), n = nbN(s, c), so = nbS(s, c);
            double dfx = 0, dfy = 0;
            if (s->Vx[c] > 0)
                dfx = (k->NUx / s->Vx[c]) *
                          (s->Nx[e] * (u[e] - u[c]) / dx - s->Nx[c] * (-u[w] + u[c]) / dx) +
                      (k->NUy / s->Vx[c]) *
                          (s->Oy[c] * (u[n] - u[c]) / dy - s->Oy[so] * (-u[so] + u[c]) / dy);
            if (s->Vy[c] > 0)
                dfy = (k->NUx / s->Vy[c]) *
                          (s->Nx[c] * (v[e] - v[c]) / dx - s->Nx[w] * (-v[w] + v[c]) / dx) +
                      (k->NUy / s->Vy[c]) *
                          (s->Oy[n] * (v[n] - v[c]) / dy - s->Oy[c] * (-v[so] + v[c]) / dy);
            s->EnXP[c] = s->EnXP[c] + dt * dfx;
            s->EnYP[c] = s->EnYP[c] + dt * dfy;
        }
    } else
    /* diffusionTerm, nsfunctions.c:589-606 */
    for (int c = 0; c < N; c++) {
        int e = nbP(s, c), w = nbM(s, c), n = nbN(s, c), so = nbS(s, c);
        double dfx = (k->NUx / (dx * dx)) * (u[e] - 2 * u[c] + u[w])
                   + (k->NUy / (dy * dy)) * (u[so] - 2 * u[c] + u[n]);
        double dfy = (k->NUx / (dx * dx)) * (v[e] - 2 * v[c] + v[w])
                   + (k->NUy / (dy * dy)) * (v[so] - 2 * v[c] + v[n]);
        s->EnXP[c] = s->EnXP[c] + dt * dfx;
        s->EnYP[c] = s->EnYP[c] + dt * dfy;
    }

    /* windTerm, nsfunctions.c:609-648 (the plain term on the subgrid path too,
       nssolve.c:131), with the reference's pi and water density */
    if (k->useWind == 1) {
        const double pi = 3.1415926, rho = 1000;
        const double phi = s->winddir[s->tt] + k->northAngle;
        const double omega = phi * pi / 180.0;
        for (int c = 0; c < N; c++) {
            double tau = k->rhoa * k->Cw *
                         (s->windspd[s->tt] - u[c] * cos(omega) - v[c] * sin(omega)) *
                         (s->windspd[s->tt] - u[c] * cos(omega) - v[c] * sin(omega));
            double tauXP = (s->depthXP[c] < k->hD && k->useThinLayer == 1)
                               ? tau * exp(k->CwT * (s->depthXP[c] - k->hD) / k->hD) : tau;
            double tauYP = (s->depthYP[c] < k->hD && k->useThinLayer == 1)
                               ? tau * exp(k->CwT * (s->depthYP[c] - k->hD) / k->hD) : tau;
            if (s->depthXP[c] > 0)
                s->EnXP[c] = s->EnXP[c] + dt * tauXP * cos(omega) / (s->depthXP[c] * rho);
            if (s->depthYP[c] > 0)
                s->EnYP[c] = s->EnYP[c] - dt * tauYP * sin(omega) / (s->depthYP[c] * rho);
        }
    }

    /* dragTerm, nsfunctions.c:651-688 / dragTermSub (staggeredV = 1),
       subgrid.c:857-927 */
    for (int c = 0; c < N; c++) {
        double velx = sqrt(u[c] * u[c] + vX[c] * vX[c]);
        double vely = sqrt(uY[c] * uY[c] + v[c] * v[c]);
        if (s->sub) {
            double effHx = s->Vx[c] > 0 ? s->Zx[c] / s->Vx[c] : 0;
            double effHy = s->Vy[c] > 0 ? s->Zy[c] / s->Vy[c] : 0;
            s->DragXP[c] = 0.5 * dt * s->CDXP[c] * velx * effHx;
            s->DragYP[c] = 0.5 * dt * s->CDYP[c] * vely * effHy;
        } else {
            s->DragXP[c] = s->depthXP[c] > 0 ? 0.5 * dt * s->CDXP[c] * velx / s->depthXP[c] : 0;
            s->DragYP[c] = s->depthYP[c] > 0 ? 0.5 * dt * s->CDYP[c] * vely / s->depthYP[c] : 0;
        }
    }
    for (int i = 0; i < s->nx; i++) s->DragXP[gjM(s, i)] = s->DragXP[i];
    for (int j = 0; j < s->ny; j++) s->DragYP[giM(s, j)] = s->DragYP[j * s->nx];

    /* dragInversion, nsfunctions.c:730-746 */
    for (int c = 0; c < N; c++) {
        s->EnXP[c] = s->EnXP[c] / (1 + s->DragXP[c]);
        s->EnYP[c] = s->EnYP[c] / (1 + s->DragYP[c]);
    }
    for (int i = 0; i < s->nx; i++) s->EnXP[gjM(s, i)] = s->EnXP[i];
    for (int j = 0; j < s->ny; j++) s->EnYP[giM(s, j)] = s->EnYP[j * s->nx];

    if (s->sub) {
        /* matrixSourceTermSub (staggeredV = 1), subgrid.c:930-975 */
        for (int c = 0; c < N; c++) {
            int w = nbM(s, c), so = nbS(s, c);
            s->z[c] = s->surf[c] * s->Z[c]
                    - dt * (s->EnXP[c] * s->Nx[c] - s->EnXP[w] * s->Nx[w])
                    - dt * (s->EnYP[c] * s->Oy[c] - s->EnYP[so] * s->Oy[so]);
        }
        if (k->bcType != 3)
            for (int q = 0; q < k->inflowLocLength; q++) {
                int c = s->inflowLoc[q];
                s->z[c] = s->z[c] + dt * (s->inflow[s->tt] / k->inflowLocLength);
            }
        for (int c = 0; c < N; c++)
            if (s->z[c] == 0 || s->z[c] == s->surf[c] * dx * dy) s->z[c] = s->surf[c] * dx * dy;
        return;
    }
    /* matrixSourceTerm, nsfunctions.c:767-797; z kept in cell order here
       (the reference stores it in the transposed solver order) */
    for (int c = 0; c < N; c++) {
        int w = nbM(s, c), so = nbS(s, c);
        s->z[c] = s->surf[c]
                - (dt / dx) * (s->EnXP[c] * s->depthXP[c] - s->EnXP[w] * s->depthXP[w])
                - (dt / dy) * (s->EnYP[c] * s->depthYP[c] - s->EnYP[so] * s->depthYP[so]);
    }
    for (int q = 0; q < k->inflowLocLength; q++) {
        int c = s->inflowLoc[q];
        s->z[c] = s->z[c] + (dt / (dx * dy)) * (s->inflow[s->tt] / k->inflowLocLength);
    }
}

/* ---- the LASpack CG (itersolv.c:355-456) in the transposed row order ------- */
typedef struct {
    int nx, ny, N;
    /* row k (k = i*ny + j, cell j*nx + i): coefficients of columns
       k-ny (W), k-1 (S), k (C), k+1 (N), k+ny (E); absent entries are flagged */
    const double *Ct, *XP, *XM, *YP, *YM;   /* cell order */
} fo_mat;

static inline int trps_cell(const fo_mat *A, int k) { return (k % A->ny) * A->nx + k / A->ny; }

/* q = A p, row-wise, entries in increasing column (operats.c:870-888) */
static void mat_vec(const fo_mat *A, const double *p, double *q)
{
    const int ny = A->ny, N = A->N;
    for (int k = 0; k < N; k++) {
        int c = trps_cell(A, k), i = k / ny, j = k % ny;
        double sum = 0.0;
        if (i > 0) sum += (-A->XM[c]) * p[k - ny];
        if (j > 0) sum += (-A->YM[c]) * p[k - 1];
        sum += A->Ct[c] * p[k];
        if (j < ny - 1) sum += (-A->YP[c]) * p[k + 1];
        if (i < A->nx - 1) sum += (-A->XP[c]) * p[k + ny];
        q[k] = sum;
    }
}

/* SSOR with omega = 1 (precond.c:48-73): forward solve with D+L
   (operats.c:1136-1160), multiply by D, backward solve with D+U
   (operats.c:1081-1105); inverse diagonal cached as 1/diag (qmatrix.c:466) */
static void ssor(const fo_mat *A, const double *invd, const double *r, double *y, double *t)
{
    const int ny = A->ny, N = A->N;
    for (int k = 0; k < N; k++) {
        int c = trps_cell(A, k), i = k / ny, j = k % ny;
        double sum = 0.0;
        if (i > 0) sum -= (-A->XM[c]) * t[k - ny];
        if (j > 0) sum -= (-A->YM[c]) * t[k - 1];
        sum += r[k];
        t[k] = sum * invd[k];
    }
    for (int k = 0; k < N; k++) {
        int c = trps_cell(A, k);
        double sum = 0.0;
        sum += A->Ct[c] * t[k];
        t[k] = sum;
    }
    for (int k = N - 1; k >= 0; k--) {
        int c = trps_cell(A, k), i = k / ny, j = k % ny;
        double sum = 0.0;
        if (i < A->nx - 1) sum -= (-A->XP[c]) * y[k + ny];
        if (j < ny - 1) sum -= (-A->YP[c]) * y[k + 1];
        sum += t[k];
        y[k] = sum * invd[k];
    }
}

static double dot(const double *a, const double *b, int n)
{
    double s = 0.0;
    for (int k = 0; k < n; k++) s += a[k] * b[k];
    return s;
}

static double l2norm(const double *a, int n)
{
    double s = 0.0;
    for (int k = 0; k < n; k++) { double c = a[k]; s += c * c; }
    return sqrt(s);
}

/* PCG from x = 0 (solveMatrix, nsfunctions.c:1004-1017 + CGIter) */
static int cg_solve(const fo_mat *A, const double *b, double *x, double eps, int maxIter,
                    int precond, double *relres)
{
    const int N = A->N;
    double *r = zalloc(N), *p = zalloc(N), *q = zalloc(N), *zz = zalloc(N), *t = zalloc(N);
    double *invd = zalloc(N);
    for (int k = 0; k < N; k++) invd[k] = 1.0 / A->Ct[trps_cell(A, k)];
    for (int k = 0; k < N; k++) x[k] = 0.0;
    double bNorm = l2norm(b, N);
    memcpy(r, b, sizeof(double) * (size_t)N);   /* x = 0: r = b (itersolv.c:395-402) */
    int iter = 0;
    double rho, rhoOld = 0.0, alpha, beta, rNorm;
    for (;;) {
        rNorm = l2norm(r, N);
        if (rNorm < eps * bNorm || (LAS_IS_ZERO(bNorm) && LAS_IS_ONE(1.0 + rNorm))) break;
        if (iter >= maxIter) break;
        iter++;
        if (precond == FREHD_PRECOND_SSOR) ssor(A, invd, r, zz, t);
        else for (int k = 0; k < N; k++) zz[k] = r[k] * invd[k];
        rho = dot(r, zz, N);
        if (iter == 1) {
            memcpy(p, zz, sizeof(double) * (size_t)N);
        } else {
            beta = rho / rhoOld;
            if (LAS_IS_ONE(beta)) for (int k = 0; k < N; k++) p[k] = zz[k] + p[k];
            else for (int k = 0; k < N; k++) p[k] = zz[k] + beta * p[k];
        }
        mat_vec(A, p, q);
        alpha = rho / dot(p, q, N);
        if (LAS_IS_ONE(alpha)) {
            for (int k = 0; k < N; k++) x[k] += p[k];
            for (int k = 0; k < N; k++) r[k] -= q[k];
        } else {
            for (int k = 0; k < N; k++) x[k] += alpha * p[k];
            for (int k = 0; k < N; k++) r[k] -= alpha * q[k];
        }
        rhoOld = rho;
    }
    if (relres) *relres = LAS_IS_ZERO(bNorm) ? 1.0 : rNorm / bNorm;
    free(r); free(p); free(q); free(zz); free(t); free(invd);
    return iter;
}

int fo_solve(int nx, int ny, const double *GnCt, const double *GnXP, const double *GnXM,
             const double *GnYP, const double *GnYM, const double *z, double *x,
             double eps, int maxIter, int precond, double *relres)
{
    fo_mat A = {nx, ny, nx * ny, GnCt, GnXP, GnXM, GnYP, GnYM};
    const int N = nx * ny;
    double *b = zalloc(N), *xt = zalloc(N);
    for (int k = 0; k < N; k++) b[k] = z[trps_cell(&A, k)];
    int it = cg_solve(&A, b, xt, eps, maxIter, precond, relres);
    for (int k = 0; k < N; k++) x[trps_cell(&A, k)] = xt[k];
    free(b); free(xt);
    return it;
}

/* ---- Phase F: solveFreeSurface (nssolve.c:158-178), plain grid ------------- */
static void phase_surface(fo_state *s)
{
    const frehd_config *k = &s->cfg;
    const int nx = s->nx, ny = s->ny, N = s->N, tt = s->tt;
    const double dt = k->dt;

    if (s->sub) {
        /* matrixCoeffSub (useSubgrid = 1, staggeredV = 1), subgrid.c:978-1081 */
        const double cx = k->g * (dt * dt), cy = k->g * (dt * dt), dxdy = k->dx * k->dy;
        const double divider = 1.0;
        for (int c = 0; c < N; c++) {
            int w = nbM(s, c), so = nbS(s, c);
            s->GnXP[c] = 0; s->GnXM[c] = 0; s->GnYP[c] = 0; s->GnYM[c] = 0;
            if (s->Vx[c] > 0) s->GnXP[c] = cx * s->Nx[c] * s->Nx[c] / s->Vx[c] / divider / (1 + s->DragXP[c]);
            if (s->Vx[w] > 0) s->GnXM[c] = cx * s->Nx[w] * s->Nx[w] / s->Vx[w] / divider / (1 + s->DragXP[w]);
            if (s->Vy[c] > 0) s->GnYP[c] = cy * s->Oy[c] * s->Oy[c] / s->Vy[c] / divider / (1 + s->DragYP[c]);
            if (s->Vy[so] > 0) s->GnYM[c] = cy * s->Oy[so] * s->Oy[so] / s->Vy[so] / divider / (1 + s->DragYP[so]);
            s->GnCt[c] = s->Z[c] + s->GnXP[c] + s->GnYP[c] + s->GnXM[c] + s->GnYM[c];
        }
        for (int c = 0; c < N; c++) {
            if (s->GnCt[c] == 0.0 || s->GnCt[c] == dxdy) s->GnCt[c] = dxdy;
            if (s->depth[c] == 0 && s->uuXP[c] == 0 && s->uuXP[nbM(s, c)] == 0 && s->vvYP[c] == 0 &&
                s->vvYP[nbS(s, c)] == 0) {
                s->GnCt[c] = dxdy;
                s->GnXM[c] = 0.0; s->GnXP[c] = 0.0; s->GnYM[c] = 0.0; s->GnYP[c] = 0.0;
                s->z[c] = s->surf[c] * dxdy;
            }
        }
    } else {
    /* matrixCoeff, nsfunctions.c:815-843 */
    double coefx = k->g * (dt * dt / (k->dx * k->dx));
    double coefy = k->g * (dt * dt / (k->dy * k->dy));
    for (int c = 0; c < N; c++) {
        int w = nbM(s, c), so = nbS(s, c);
        s->GnXP[c] = coefx * s->depthXP[c] / (1 + s->DragXP[c]);
        s->GnXM[c] = coefx * s->depthXP[w] / (1 + s->DragXP[w]);
        s->GnYP[c] = coefy * s->depthYP[c] / (1 + s->DragYP[c]);
        s->GnYM[c] = coefy * s->depthYP[so] / (1 + s->DragYP[so]);
        s->GnCt[c] = 1 + s->GnXP[c] + s->GnYP[c] + s->GnXM[c] + s->GnYM[c];
    }
    }
    /* adjustMatrixBoundary, nsfunctions.c:846-884: Neumann walls on all sides */
    for (int j = 0; j < ny; j++) {
        int a = j * nx, b = j * nx + nx - 1;
        s->GnCt[a] = s->GnCt[a] - s->GnXM[a]; s->GnXM[a] = 0;
        s->GnCt[b] = s->GnCt[b] - s->GnXP[b]; s->GnXP[b] = 0;
    }
    for (int i = 0; i < nx; i++) {
        int a = i, b = N - nx + i;
        s->GnCt[a] = s->GnCt[a] - s->GnYM[a]; s->GnYM[a] = 0;
        s->GnCt[b] = s->GnCt[b] - s->GnYP[b]; s->GnYP[b] = 0;
    }
    /* adjustTidalBoundary, nsfunctions.c:887-924: Dirichlet rows */
    if (k->bcType != 2)
        for (int q = 0; q < k->tideLocLengthP; q++) {
            int c = s->tideLocP[q];
            s->GnCt[c] = 1; s->GnXM[c] = 0; s->GnXP[c] = 0; s->GnYM[c] = 0; s->GnYP[c] = 0;
            s->z[c] = s->tideP[tt] + k->offset;
        }
    if (k->bcType != 1)
        for (int q = 0; q < k->tideLocLengthM; q++) {
            int c = s->tideLocM[q];
            s->GnCt[c] = 1; s->GnXM[c] = 0; s->GnXP[c] = 0; s->GnYM[c] = 0; s->GnYP[c] = 0;
            s->z[c] = s->tideM[tt] + k->offset;
        }

    /* setupMatrix + solveMatrix + getFreeSurface, nsfunctions.c:927-1031 */
    double *x = zalloc(N);
    s->pcgIters = fo_solve(nx, ny, s->GnCt, s->GnXP, s->GnXM, s->GnYP, s->GnYM, s->z, x,
                           k->eps, k->maxIter, k->precond, &s->relres);
    s->pcgItersTotal += s->pcgIters;
    memcpy(s->surfOld, s->surf, sizeof(double) * (size_t)s->Ng);
    for (int c = 0; c < N; c++) s->surf[c] = x[c];
    free(x);

    zero_neg_surf(s, tt);
    fill_surface_ghosts(s, s->surf);
    tidal_surface(s, tt);
}

/* ---- Phase U: updateData (nssolve.c:181-290), plain grid -------------------- */
static void phase_update(fo_state *s)
{
    const frehd_config *k = &s->cfg;
    const int nx = s->nx, ny = s->ny, N = s->N, Ng = s->Ng, tt = s->tt;
    const double dx = k->dx, dy = k->dy, dt = k->dt;
    double *h = s->surf, *ho = s->surfOld, *zb = s->bottomZ, *d = s->depth;

    /* limiterCFL, nsfunctions.c:452-495 */
    for (int c = 0; c < N; c++) {
        double diff = h[c] - zb[c];
        if (d[c] <= 0 && diff > 0) {
            double sref = ho[c] + k->minDepth;
            int e = nbP(s, c), w = nbM(s, c), n = nbN(s, c), so = nbS(s, c);
            /* depth holds interior values only: a ghost neighbour reads 0 */
            int dE = ho[e] <= sref || (e < N ? d[e] : 0.0) <= 0;
            int dW = ho[w] <= sref || (w < N ? d[w] : 0.0) <= 0;
            int dN = ho[n] <= sref || (n < N ? d[n] : 0.0) <= 0;
            int dS = ho[so] <= sref || (so < N ? d[so] : 0.0) <= 0;
            if (dE && dW && dN && dS) {
                if ((double)c < (double)N - dx) s->Vloss[0] += diff * dx * dy;
                h[c] = zb[c];
                s->rmvCFL[c] = 1;
            }
        }
    }
    for (int c = 0; c < N; c++) {
        double diff = h[c] - zb[c];
        if (diff <= k->minDepth && diff > 0) {
            s->Vloss[1] += diff * dx * dy;
            h[c] = zb[c];
            s->cellV[c] = 0;
            s->rmvCFL[c] = 0;
        }
    }

    if (s->sub) {
        /* computeVolumeSub (subgrid.c:1215-1224): the previous step's V */
        for (int c = 0; c < N; c++) {
            s->cellV[c] = s->V[c];
            if (k->useScalar == 1)
                for (int q = 0; q < k->nScalar; q++)
                    s->Sm[q * Ng + c] = s->S[q * Ng + c] * s->cellV[c];
        }
        compute_sub_area(s);
        combine_sub_area(s);
    } else {
        compute_volume(s);
    }
    update_depth(s);
    update_face_depth(s);

    double *u = s->uuXP, *v = s->vvYP;
    if (s->sub) {
        /* detectWaterfallLocationSub, subgrid.c:1335-1397: flags reset here,
           depth test on the neighbour's surface above the face bottom */
        for (int c = 0; c < N; c++) { s->wtfXP[c] = 0; s->wtfYP[c] = 0; }
        for (int c = 0; c < N; c++) {
            double bot = s->bottomXP[c];
            if (h[c] < bot - 0.0 && h[nbP(s, c)] - bot >= k->wtfh) s->wtfXP[c] = -1;
        }
        for (int c = 0; c < N; c++) {
            int w = nbM(s, c);
            double bot = s->bottomXP[w];
            if (h[c] < bot - 0.0 && h[w] - bot >= k->wtfh) s->wtfXP[w] = 1;
        }
        for (int c = 0; c < N; c++) {
            double bot = s->bottomYP[c];
            if (h[c] < bot - 0.0 && h[nbN(s, c)] - bot >= k->wtfh) s->wtfYP[c] = -1;
        }
        for (int c = 0; c < N; c++) {
            int so = nbS(s, c);
            double bot = s->bottomYP[so];
            if (h[c] < bot - 0.0 && h[so] - bot >= k->wtfh) s->wtfYP[so] = 1;
        }
        /* updateVelocitySub (useSubgrid = 1, staggeredV = 1), subgrid.c:1084-1187 */
        const double divider = 1.0;
        for (int c = 0; c < N; c++) {
            double effHx = s->Vx[c] > 0 ? s->Nx[c] / s->Vx[c] : 0;
            double effHy = s->Vy[c] > 0 ? s->Oy[c] / s->Vy[c] : 0;
            u[c] = s->EnXP[c] - (k->g * dt * effHx) * (h[nbP(s, c)] - h[c]) / divider;
            v[c] = s->EnYP[c] - (k->g * dt * effHy) * (h[nbN(s, c)] - h[c]) / divider;
        }
        for (int c = 0; c < N; c++) {
            if (s->Nx[c] == 0) u[c] = 0;
            if (s->Oy[c] == 0) v[c] = 0;
        }
        /* no outflow from a dry cell (in place, neighbour faces included) */
        for (int c = 0; c < N; c++) {
            if (d[c] == 0) {
                int w = nbM(s, c), so = nbS(s, c);
                if (u[c] > 0) u[c] = 0;
                if (u[w] < 0) u[w] = 0;
                if (v[c] > 0) v[c] = 0;
                if (v[so] < 0) v[so] = 0;
            }
        }
        for (int c = 0; c < N; c++) {
            if (s->rmvCFL[c] == 1) {
                u[c] = 0; v[c] = 0; u[nbM(s, c)] = 0; v[nbS(s, c)] = 0;
                s->rmvCFL[c] = 0;
            }
        }
        /* waterfallCorrectionSub, subgrid.c:1400-1444 (flags kept) */
        const double CFLmax = 0.24 / k->subR, Cw = 0.7;
        for (int c = 0; c < N; c++) {
            if (s->wtfXP[c] != 0) {
                double vel0 = u[c];
                u[c] = s->wtfXP[c] * Cw * sqrt(2 * k->g * s->Nx[c] / dy);
                if (fabs(u[c]) > CFLmax * dx / dt) u[c] = s->wtfXP[c] * CFLmax * dx / dt;
                s->Vloss[3] -= (vel0 - u[c]) * dt * s->Nx[c];
            }
            if (s->wtfYP[c] != 0) {
                double vel0 = v[c];
                v[c] = s->wtfYP[c] * Cw * sqrt(2 * k->g * s->Oy[c] / dx);
                if (fabs(v[c]) > CFLmax * dy / dt) v[c] = s->wtfYP[c] * CFLmax * dy / dt;
                s->Vloss[3] -= (vel0 - v[c]) * dt * s->Oy[c];
            }
        }
    } else {
    /* detectWaterfallLocation, nsfunctions.c:1080-1130 (dmin = 0) */
    for (int c = 0; c < N; c++)
        if (h[c] < s->bottomXP[c] - 0.0 && s->depthXP[c] >= k->wtfh) s->wtfXP[c] = -1;
    for (int c = 0; c < N; c++) {
        int w = nbM(s, c);
        if (h[c] < s->bottomXP[w] - 0.0 && s->depthXP[w] >= k->wtfh) s->wtfXP[w] = 1;
    }
    for (int c = 0; c < N; c++)
        if (h[c] < s->bottomYP[c] - 0.0 && s->depthYP[c] >= k->wtfh) s->wtfYP[c] = -1;
    for (int c = 0; c < N; c++) {
        int so = nbS(s, c);
        if (h[c] < s->bottomYP[so] - 0.0 && s->depthYP[so] >= k->wtfh) s->wtfYP[so] = 1;
    }

    /* updateVelocity, nsfunctions.c:250-286: not divided by (1 + Drag) */
    for (int c = 0; c < N; c++) {
        u[c] = s->EnXP[c] - (k->g * dt / dx) * (h[nbP(s, c)] - h[c]);
        v[c] = s->EnYP[c] - (k->g * dt / dy) * (h[nbN(s, c)] - h[c]);
    }
    for (int c = 0; c < N; c++) {
        if (s->depthXP[c] == 0) u[c] = 0;
        if (s->depthYP[c] == 0) v[c] = 0;
    }
    for (int c = 0; c < N; c++) {
        if (s->rmvCFL[c] == 1) {
            u[c] = 0; v[c] = 0; u[nbM(s, c)] = 0; v[nbS(s, c)] = 0;
            s->rmvCFL[c] = 0;
        }
    }

    /* waterfallCorrection, nsfunctions.c:1133-1166 (Cw = 0.7, CFLmax = 0.24) */
    for (int c = 0; c < N; c++) {
        if (s->wtfXP[c] != 0) {
            double vel0 = u[c];
            u[c] = s->wtfXP[c] * 0.7 * sqrt(2 * k->g * s->depthXP[c]);
            if (u[c] > 0.24 * dx / dt) u[c] = 0.24 * dx / dt;
            s->Vloss[3] -= (vel0 - u[c]) * dt * s->depthXP[c] * dy;
        }
        if (s->wtfYP[c] != 0) {
            double vel0 = v[c];
            v[c] = s->wtfYP[c] * 0.7 * sqrt(2 * k->g * s->depthYP[c]);
            if (v[c] > 0.24 * dy / dt) v[c] = 0.24 * dy / dt;
            s->Vloss[3] -= (vel0 - v[c]) * dt * s->depthYP[c] * dx;
        }
    }
    for (int c = 0; c < N; c++) { s->wtfXP[c] = 0; s->wtfYP[c] = 0; }
    }

    fill_velocity_ghosts(s);

    /* computeFaceFlowRate[Sub], nsfunctions.c:350-358 / subgrid.c:1190-1212 (over Ng) */
    for (int c = 0; c < Ng; c++) {
        if (s->sub) {
            s->Fuu[c] = u[c] * s->Nx[c];
            s->Fvv[c] = v[c] * s->Oy[c];
        } else {
            s->Fuu[c] = u[c] * dy * s->depthXP[c];
            s->Fvv[c] = v[c] * dx * s->depthYP[c];
        }
    }

    /* adjustTidalVelocity[Sub], nsfunctions.c:1034-1078 / subgrid.c:1291-1332 */
    if (k->bcType != 2)
        for (int q = 0; q < k->tideLocLengthP; q++) {
            int c = s->tideLocP[q];
            if (s->sub) {
                double Fww = (h[c] - ho[c]) * s->Z[c] / dt;
                double all = s->Fvv[nbS(s, c)] + s->Fuu[nbM(s, c)] - s->Fuu[c] - Fww;
                if (all != 0 && s->Oy[c] > 0) v[c] = all / s->Oy[c];
                else v[c] = 0;
            } else {
                double Fww = (h[c] - ho[c]) * dx * dy / dt;
                double all = s->Fvv[nbS(s, c)] + s->Fuu[nbM(s, c)] - s->Fuu[c] - Fww;
                if (all != 0 && d[c] > 0) v[c] = all / (d[c] * dx);
                else v[c] = 0;
            }
        }
    if (k->bcType != 1)
        for (int q = 0; q < k->tideLocLengthM; q++) {
            int c = s->tideLocM[q];
            if (s->sub) {
                double Fww = (h[c] - ho[c]) * s->Z[c] / dt;
                double all = -s->Fvv[c] + s->Fuu[nbM(s, c)] - s->Fuu[c] - Fww;
                if (all != 0 && s->Oy[c] > 0) v[nbS(s, c)] = -all / s->Oy[c];
                else v[nbS(s, c)] = 0;
            } else {
                double Fww = (h[c] - ho[c]) * dx * dy / dt;
                double all = -s->Fvv[c] + s->Fuu[nbM(s, c)] - s->Fuu[c] - Fww;
                if (all != 0 && d[c] > 0) v[nbS(s, c)] = -all / (d[c] * dx);
            }
        }

    velocity_interp(s);

    /* volumeByFlux[Sub], nsfunctions.c:373-409 / subgrid.c:1227-1263 */
    const int inflowOn = s->sub ? (k->bcType != 3) : 1;
    for (int c = 0; c < N; c++) {
        s->cellV[c] = s->cellV[c] + dt * (s->Fuu[nbM(s, c)] - s->Fuu[c] + s->Fvv[nbS(s, c)] - s->Fvv[c]);
        if (inflowOn)
            for (int q = 0; q < k->inflowLocLength; q++)
                if (c == s->inflowLoc[q]) s->cellV[c] = s->cellV[c] + s->inflow[tt] * dt;
        if (s->cellV[c] < 0) s->cellV[c] = 0;
    }

    /* updateCD, nsfunctions.c:1170-1189 / updateCDSub, subgrid.c:1446-1535:
       useSubDrag 2 (Volp 2013, :1481-1506), 3 and 4 (Casas 2010, :1508-1519),
       otherwise Manning with h = V/Z (:1520-1535) */
    if (k->CDnotN == 0 && s->sub && k->useSubDrag == 2) {
        for (int c = 0; c < N; c++) {
            double coeff = manning_coeff(s, c);
            if (s->YhX[c] > 0) {
                s->CDXP[c] = coeff * s->YhX[c];
                if (s->CDXP[c] > k->CDmax) s->CDXP[c] = k->CDmax;
            } else {
                s->CDXP[c] = 0;
            }
            if (s->YhY[c] > 0) {
                s->CDYP[c] = coeff * s->YhY[c];
                if (s->CDYP[c] > k->CDmax) s->CDYP[c] = k->CDmax;
            } else {
                s->CDYP[c] = 0;
            }
        }
    } else if (k->CDnotN == 0 && s->sub && (k->useSubDrag == 4 || k->useSubDrag == 3)) {
        for (int c = 0; c < N; c++) {
            if (s->V[c] > 0) {
                double h = s->V[c] / (k->dx * k->dy);
                s->CDXP[c] = k->g * s->Cn[c] * s->Cn[c] / pow(h, 1.0 / 3.0);
                s->CDYP[c] = k->g * s->Cn[c] * s->Cn[c] / pow(h, 1.0 / 3.0);
            }
        }
    } else if (k->CDnotN == 0) {
        for (int c = 0; c < N; c++) {
            if (s->sub) {
                if (s->V[c] > 0) {
                    double coeff = manning_coeff(s, c);
                    double hh = s->V[c] / s->Z[c];
                    s->CDXP[c] = coeff / pow(hh, 1.0 / 3.0);
                    s->CDYP[c] = coeff / pow(hh, 1.0 / 3.0);
                }
            } else if (d[c] > 0) {
                double coeff = manning_coeff(s, c);
                s->CDXP[c] = coeff / pow(d[c], 1.0 / 3.0);
                s->CDYP[c] = coeff / pow(d[c], 1.0 / 3.0);
            }
        }
    }
    if (k->useThinLayer == 1) thin_layer_drag(s);
    (void)ny; (void)nx;
}

/* ---- Phase T: scalarTransport (scalar.c:30-56), upwind, plain grid ---------- */
static int is_inflow_cell(const fo_state *s, int c)
{
    for (int q = 0; q < s->cfg.inflowLocLength; q++) if (s->inflowLoc[q] == c) return 1;
    return 0;
}

/* superbee limiter of the TVD scheme (scalar.c:113-117) */
static double superbee(double r)
{
    double p1, p2, phi;
    if (2 * r > 1) p1 = 1; else p1 = 2 * r;
    if (r > 2) p2 = 2; else p2 = r;
    if (p1 > 0) phi = p1; else phi = 0;
    if (p2 > phi) phi = p2;
    return phi;
}

/* face value of the TVD scheme (scalar.c:103-245) for the face between cell
   L and the next cell H (x or y): F the flow rate at the face, FL the one
   at the face on L's far side, FH at H's far side; SLL, SHH the cells one
   further out; Cu the face's Courant number.  Upwind from L when F > 0 (first
   order if the flow at L's far side is not also positive), else from H. */
static double tvd_face(double SL, double SH, double SLL, double SHH, double F, double FL, double FH, double Cu)
{
    const double dSmin = 0.00001;
    double r;
    if (F > 0) {
        if (FL <= 0) return SL;
        if (fabs(SH - SL) < dSmin) r = 0;
        else r = (SL - SLL) / (SH - SL);
        return SL + 0.5 * superbee(r) * (SH - SL) * Cu * (1 - Cu);
    }
    if (FH >= 0) return SH;
    if (fabs(SH - SL) < dSmin) r = 0;
    else r = (SH - SHH) / (SL - SH);
    return SH + 0.5 * superbee(r) * (SL - SH) * Cu * (1 - Cu);
}

static void scalar_pass(fo_state *s, double *S, double *Sm)
{
    const frehd_config *k = &s->cfg;
    const int nx = s->nx, ny = s->ny, N = s->N, tt = s->tt;
    const double dx = k->dx, dy = k->dy, dt = k->dt;
    const double *Fu = s->Fuu, *Fv = s->Fvv, *ho = s->surfOld, *zb = s->bottomZ;

    /* scalarAdvection, scalar.c:68-97 (upwind) and :99-255 (TVD superbee) */
    for (int c = 0; c < N; c++) {
        int e = nbP(s, c), w = nbM(s, c), n = nbN(s, c), so = nbS(s, c);
        double Sd, Su, Sr, Sl;
        if (k->scalarAdv == 2) {
            const double *u = s->uuXP, *v = s->vvYP;
            Sd = tvd_face(S[c], S[e], S[w], S[nbPP(s, c)], Fu[c], Fu[w], Fu[e], u[c] * dt / dx);
            Su = tvd_face(S[w], S[c], S[nbMM(s, c)], S[e], Fu[w], Fu[nbMM(s, c)], Fu[c], u[w] * dt / dx);
            Sr = tvd_face(S[c], S[n], S[so], S[nbNN(s, c)], Fv[c], Fv[so], Fv[n], v[c] * dt / dy);
            /* the YM face's Courant number divides by dx (scalar.c:222) */
            Sl = tvd_face(S[so], S[c], S[nbSS(s, c)], S[n], Fv[so], Fv[nbSS(s, c)], Fv[c], v[so] * dt / dx);
        } else {
            Sd = Fu[c] > 0 ? S[c] : S[e];
            Su = Fu[w] > 0 ? S[w] : S[c];
            Sr = Fv[c] > 0 ? S[c] : S[n];
            Sl = Fv[so] > 0 ? S[so] : S[c];
        }
        Sm[c] = Sm[c] + dt * (-Fu[c] * Sd + Fu[w] * Su - Fv[c] * Sr + Fv[so] * Sl);
    }
    if (s->sub) {
        /* scalarDiffusionSub, subgrid.c:1552-1567 */
        for (int c = 0; c < N; c++) {
            int e = nbP(s, c), w = nbM(s, c), n = nbN(s, c), so = nbS(s, c);
            Sm[c] = Sm[c] + dt *
                ((k->Kx * s->Nx[c] / dx) * (S[e] - S[c]) -
                 (k->Kx * s->Nx[w] / dx) * (S[c] - S[w]) +
                 (k->Ky * s->Oy[c] / dy) * (S[n] - S[c]) -
                 (k->Ky * s->Oy[so] / dy) * (S[c] - S[so]));
        }
    } else
    /* scalarDiffusion, scalar.c:259-274 */
    for (int c = 0; c < N; c++) {
        int e = nbP(s, c), w = nbM(s, c), n = nbN(s, c), so = nbS(s, c);
        Sm[c] = Sm[c] + dt *
            ((k->Kx * dy * s->depthXP[c] / dx) * (S[e] - S[c]) -
             (k->Kx * dy * s->depthXP[w] / dx) * (S[c] - S[w]) +
             (k->Ky * dx * s->depthYP[c] / dy) * (S[n] - S[c]) -
             (k->Ky * dx * s->depthYP[so] / dy) * (S[c] - S[so]));
    }
    /* updateScalar, scalar.c:277-416 */
    int *cnt = (int *)calloc((size_t)N, sizeof(int));
    double *smax = zalloc(N), *smin = zalloc(N);
    for (int c = 0; c < N; c++) {
        int e = nbP(s, c), w = nbM(s, c), n = nbN(s, c), so = nbS(s, c);
        double vals[4]; int m = 0;
        /* faces open: depth (plain) or subgrid face area (updateScalarSub,
           subgrid.c:1588-1630) */
        const double *fx = s->sub ? s->Nx : s->depthXP, *fy = s->sub ? s->Oy : s->depthYP;
        if (fx[c] > 0 && ho[e] - zb[e] > 0) vals[m++] = S[e];
        if (fx[w] > 0 && ho[w] - zb[w] > 0) vals[m++] = S[w];
        /* last strip: only neighbour indices <= N count (scalar.c:308-315) */
        if (n <= N && fy[c] > 0 && ho[n] - zb[n] > 0) vals[m++] = S[n];
        if (so <= N && fy[so] > 0 && ho[so] - zb[so] > 0) vals[m++] = S[so];
        cnt[c] = m;
        if (m) {
            double mx = vals[0], mn = vals[0];
            for (int q = 1; q < m; q++) { if (vals[q] > mx) mx = vals[q]; if (vals[q] < mn) mn = vals[q]; }
            smax[c] = mx; smin[c] = mn;
        }
    }
    const int inflowOn = s->sub ? (k->bcType != 3) : 1;
    if (inflowOn)
        for (int q = 0; q < k->inflowLocLength; q++) {
            int c = s->inflowLoc[q];
            Sm[c] = Sm[c] + k->inflowS * s->inflow[tt] * dt / k->inflowLocLength;
        }
    for (int c = 0; c < N; c++) {
        S[c] = s->cellV[c] > 0 ? Sm[c] / s->cellV[c] : 0;
        int fl = inflowOn && is_inflow_cell(s, c);
        if (cnt[c] != 0 && S[c] > smax[c] && !fl) S[c] = smax[c];
        else if (cnt[c] != 0 && S[c] < smin[c] && !fl) S[c] = smin[c];
    }
    free(cnt); free(smax); free(smin);
    for (int c = 0; c < N; c++) {
        if (S[c] > 100 || S[c] < 0) {
            int i = c % nx, j = c / nx;
            int bottomRow = (j == 0);                  /* icjM > N2ci */
            int topRowNotFirst = (j == ny - 1 && i > 0); /* icjP > N2ci */
            if (!bottomRow && !topRowNotFirst) {
                S[c] = 0;
                s->surf[c] = s->surf[c] - s->depth[c];
                s->depth[c] = 0;
            }
        }
    }
    /* enforceScalarBC, scalar.c:419-471 */
    for (int i = 0; i < nx; i++) { S[gjP(s, i)] = S[N - nx + i]; S[gjM(s, i)] = S[i]; }
    for (int j = 0; j < ny; j++) { S[giP(s, j)] = S[j * nx + nx - 1]; S[giM(s, j)] = S[j * nx]; }
    if (k->bcType != 2)
        for (int q = 0; q < k->tideLocLengthP; q++)
            S[s->tideLocP[q]] = k->useConstTidePS == 1 ? k->tidePS : s->tidalPS[tt];
    if (k->bcType != 1)
        for (int q = 0; q < k->tideLocLengthM; q++)
            S[s->tideLocM[q]] = k->useConstTideMS == 1 ? k->tideMS : s->tidalMS[tt];
    for (int c = 0; c < N; c++) if (s->depth[c] <= 0) S[c] = 0;
}

static void phase_scalar(fo_state *s)
{
    if (s->cfg.useScalar != 1) return;
    for (int q = 0; q < s->cfg.nScalar; q++)
        scalar_pass(s, s->S + (size_t)q * s->Ng, s->Sm + (size_t)q * s->Ng);
}

/* ---- monitorCFL, nsfunctions.c:1192-1232 ------------------------------------- */
static int phase_cfl(fo_state *s)
{
    double mx = 0, my = 0;
    for (int c = 0; c < s->N; c++) {
        s->CFLx[c] = fabs(s->uuXP[c]) * s->cfg.dt / s->cfg.dx;
        s->CFLy[c] = fabs(s->vvYP[c]) * s->cfg.dt / s->cfg.dy;
        if (s->CFLx[c] > mx) mx = s->CFLx[c];
        if (s->CFLy[c] > my) my = s->CFLy[c];
    }
    return (fabs(mx) > 20 || fabs(my) > 20) ? FREHD_ECFL : 0;
}

int fo_phase(fo_state *s, int phase)
{
    switch (phase) {
    case FO_PHASE_SOURCE:
        if (s->tt + 1 > s->cfg.Nt) return FREHD_EINVAL;   /* forcing series exhausted */
        s->tt++; phase_source(s); return 0;
    case FO_PHASE_SURFACE: phase_surface(s); return 0;
    case FO_PHASE_UPDATE: phase_update(s); return 0;
    case FO_PHASE_SCALAR: phase_scalar(s); return 0;
    case FO_PHASE_CFL: return phase_cfl(s);
    default: return FREHD_EINVAL;
    }
}

int fo_step(fo_state *s)
{
    if (fo_phase(s, FO_PHASE_SOURCE)) return FREHD_EINVAL;
    for (int p = FO_PHASE_SURFACE; p < FO_PHASE_CFL; p++) fo_phase(s, p);
    return fo_phase(s, FO_PHASE_CFL);
}

/* ---- creation: the Init sequence, initialize.c:40-155 ----------------------- */
static void *dup_arr(const void *src, size_t bytes)
{
    if (!src || bytes == 0) return NULL;
    void *d = malloc(bytes);
    memcpy(d, src, bytes);
    return d;
}

fo_state *fo_create(const frehd_config *cfg, const frehd_forcing *f, const frehd_ic *ic)
{
    if (!cfg || !f || !ic || !ic->bottomZ || cfg->nx < 4 || cfg->ny < 4) return NULL;
    fo_state *s = (fo_state *)calloc(1, sizeof(fo_state));
    s->cfg = *cfg;
    if (s->cfg.nScalar < 1) s->cfg.nScalar = 1;
    const int nx = cfg->nx, ny = cfg->ny, N = nx * ny, Ng = (nx + 2) * (ny + 2);
    const int ns = s->cfg.nScalar;
    s->nx = nx; s->ny = ny; s->N = N; s->Ng = Ng; s->tt = 0;
    size_t ser = sizeof(double) * (size_t)(cfg->Nt + 1);
    s->tideLocP = (int32_t *)dup_arr(f->tideLocP, sizeof(int32_t) * (size_t)cfg->tideLocLengthP);
    s->tideLocM = (int32_t *)dup_arr(f->tideLocM, sizeof(int32_t) * (size_t)cfg->tideLocLengthM);
    s->inflowLoc = (int32_t *)dup_arr(f->inflowLoc, sizeof(int32_t) * (size_t)cfg->inflowLocLength);
    s->tideP = (double *)dup_arr(f->tideP, ser);
    s->tideM = (double *)dup_arr(f->tideM, ser);
    s->inflow = (double *)dup_arr(f->inflow, ser);
    s->tidalPS = (double *)dup_arr(f->tidalPS, ser);
    s->tidalMS = (double *)dup_arr(f->tidalMS, ser);
    s->manningCell = (double *)dup_arr(f->manningCell, sizeof(double) * (size_t)N);
    s->windspd = (double *)dup_arr(cfg->useWind == 1 ? f->windspd : NULL, ser);
    s->winddir = (double *)dup_arr(cfg->useWind == 1 ? f->winddir : NULL, ser);
    if (!s->windspd) s->windspd = zalloc(cfg->Nt + 1);
    if (!s->winddir) s->winddir = zalloc(cfg->Nt + 1);
    if (!s->inflow) s->inflow = zalloc(cfg->Nt + 1);
    if (!s->tideP) s->tideP = zalloc(cfg->Nt + 1);
    if (!s->tideM) s->tideM = zalloc(cfg->Nt + 1);
    if (!s->tidalPS) s->tidalPS = zalloc(cfg->Nt + 1);
    if (!s->tidalMS) s->tidalMS = zalloc(cfg->Nt + 1);

    double **ngArrays[] = {&s->surf, &s->surfOld, &s->uuXP, &s->vvYP, &s->Fuu, &s->Fvv,
        &s->depthXP, &s->depthYP, &s->EnXP, &s->EnYP, &s->DragXP, &s->DragYP, &s->CDXP,
        &s->CDYP, &s->wtfXP, &s->wtfYP, &s->CFLx, &s->CFLy, &s->advX, &s->advY,
        &s->bottomZ, &s->bottomXP, &s->bottomYP, &s->uuYP, &s->vvXP, &s->depth,
        &s->cellV, &s->rmvCFL, &s->GnCt, &s->GnXP, &s->GnXM, &s->GnYP, &s->GnYM, &s->z};
    for (size_t q = 0; q < sizeof(ngArrays) / sizeof(ngArrays[0]); q++) *ngArrays[q] = zalloc(Ng);
    s->S = zalloc(ns * Ng);
    s->Sm = zalloc(ns * Ng);

    /* InitBathymetry (bathymetry.c:74-104): interior bed (and face edges for
       the subgrid) + offset */
    for (int c = 0; c < N; c++) s->bottomZ[c] = ic->bottomZ[c] + cfg->offset;
    if (cfg->useSubgrid) {
        const frehd_subgrid *sg = ic->sub;
        if (!sg || !sg->edgeX || !sg->edgeY || cfg->Ne < 1) { fo_destroy(s); return NULL; }
        s->sub = 1;
        s->Ne = cfg->Ne;
        s->edgeX = zalloc(Ng);
        s->edgeY = zalloc(Ng);
        for (int c = 0; c < N; c++) {
            s->edgeX[c] = sg->edgeX[c] + cfg->offset;
            s->edgeY[c] = sg->edgeY[c] + cfg->offset;
        }
    }
    init_bathymetry(s);
    /* initCD (initialize.c:381-389) */
    for (int c = 0; c < Ng; c++) { s->CDXP[c] = cfg->CDx; s->CDYP[c] = cfg->CDy; }
    /* initFieldValues (initialize.c:392-440) */
    for (int c = 0; c < N; c++) {
        s->surf[c] = (ic->surf0cell ? ic->surf0cell[c] : ic->surf0) + cfg->offset;
        s->uuXP[c] = ic->U0cell ? ic->U0cell[c] : ic->U0;
        s->vvYP[c] = ic->V0cell ? ic->V0cell[c] : ic->V0;
        if (cfg->useScalar == 1)
            for (int q = 0; q < ns; q++)
                s->S[(size_t)q * Ng + c] = ic->S0cell ? ic->S0cell[(size_t)q * N + c] : ic->S0;
    }
    zero_neg_surf(s, 0);
    fill_surface_ghosts(s, s->surf);
    fill_velocity_ghosts(s);
    tidal_surface(s, 0);
    update_depth(s);
    update_face_depth(s);
    if (s->sub) {
        /* initAllSubVar (subgrid.c:2112-2284), computeSubArea, combineSubArea */
        const frehd_subgrid *sg = ic->sub;
        const int Ne = s->Ne;
        const size_t tn = (size_t)Ne * N;
        s->allSurf = zalloc(Ne);
        for (int q = 0; q < Ne; q++) s->allSurf[q] = cfg->surfmin + q * cfg->dsurf + cfg->offset;
        double **tabs[] = {&s->tV, &s->tZ, &s->tVxp, &s->tVxm, &s->tVyp, &s->tVym, &s->tNp,
                           &s->tNm, &s->tOp, &s->tOm, &s->tminVxp, &s->tminVxm, &s->tminVyp,
                           &s->tminVym};
        const double *src[] = {sg->V, sg->Z, sg->Vxp, sg->Vxm, sg->Vyp, sg->Vym, sg->Np, sg->Nm,
                               sg->Op, sg->Om, sg->minVxp, sg->minVxm, sg->minVyp, sg->minVym};
        for (int q = 0; q < 14; q++) {
            *tabs[q] = (double *)calloc(tn, sizeof(double));
            if (src[q]) memcpy(*tabs[q], src[q], sizeof(double) * tn);
        }
        s->tYhX = (double *)dup_arr(cfg->useSubDrag == 2 ? sg->YhX : NULL, sizeof(double) * tn);
        s->tYhY = (double *)dup_arr(cfg->useSubDrag == 2 ? sg->YhY : NULL, sizeof(double) * tn);
        s->tCn = (double *)dup_arr(cfg->useSubDrag >= 3 ? sg->Cn : NULL, sizeof(double) * tn);
        s->tNmin = (double *)dup_arr(cfg->useminA == 1 ? sg->Nmin : NULL, sizeof(double) * tn);
        s->tNmax = (double *)dup_arr(cfg->useminA == 1 ? sg->Nmax : NULL, sizeof(double) * tn);
        s->tOmin = (double *)dup_arr(cfg->useminA == 1 ? sg->Omin : NULL, sizeof(double) * tn);
        s->tOmax = (double *)dup_arr(cfg->useminA == 1 ? sg->Omax : NULL, sizeof(double) * tn);
        s->YhX = zalloc(Ng);
        s->YhY = zalloc(Ng);
        s->Cn = zalloc(Ng);
        double **subArr[] = {&s->V, &s->Z, &s->Vxp, &s->Vyp, &s->Vxm, &s->Vym, &s->Np, &s->Nm,
                             &s->Op, &s->Om, &s->Vx, &s->Vy, &s->Zx, &s->Zy, &s->Nx, &s->Oy,
                             &s->minVxp, &s->minVxm, &s->minVyp, &s->minVym, &s->minVx,
                             &s->minVy, &s->Nmin, &s->Omin, &s->Nmax, &s->Omax};
        for (size_t q = 0; q < sizeof(subArr) / sizeof(subArr[0]); q++) *subArr[q] = zalloc(Ng);
        s->ind = (int *)calloc((size_t)Ng, sizeof(int));
        for (int c = 0; c < Ng; c++) s->ind[c] = Ne / 2;   /* round(Ne / 2), integer division */
        compute_sub_area(s);
        combine_sub_area(s);
    }
    compute_volume(s);    /* the plain computeVolume, also with the subgrid (initialize.c:134) */
    velocity_interp(s);
    thin_layer_drag(s);   /* called unconditionally by Init (initialize.c:138) */
    if (cfg->useScalar == 1) {
        for (int q = 0; q < ns; q++) {
            double *S = s->S + (size_t)q * Ng, *Sm = s->Sm + (size_t)q * Ng;
            for (int c = 0; c < N; c++) if (s->depth[c] <= 0) S[c] = 0;
            for (int c = 0; c < N; c++) Sm[c] = S[c] * s->cellV[c];   /* scalarMass */
        }
    }
    return s;
}

void fo_destroy(fo_state *s)
{
    if (!s) return;
    double *arrs[] = {s->surf, s->surfOld, s->uuXP, s->vvYP, s->Fuu, s->Fvv, s->depthXP,
        s->depthYP, s->EnXP, s->EnYP, s->DragXP, s->DragYP, s->CDXP, s->CDYP, s->wtfXP,
        s->wtfYP, s->CFLx, s->CFLy, s->advX, s->advY, s->bottomZ, s->bottomXP, s->bottomYP,
        s->uuYP, s->vvXP, s->depth, s->cellV, s->rmvCFL, s->GnCt, s->GnXP, s->GnXM, s->GnYP,
        s->GnYM, s->z, s->S, s->Sm, s->tideP, s->tideM, s->inflow, s->tidalPS, s->tidalMS,
        s->manningCell, s->windspd, s->winddir};
    for (size_t q = 0; q < sizeof(arrs) / sizeof(arrs[0]); q++) free(arrs[q]);
    double *subs[] = {s->allSurf, s->edgeX, s->edgeY, s->tV, s->tZ, s->tVxp, s->tVxm, s->tVyp,
        s->tVym, s->tNp, s->tNm, s->tOp, s->tOm, s->tminVxp, s->tminVxm, s->tminVyp, s->tminVym,
        s->V, s->Z, s->Vxp, s->Vyp, s->Vxm, s->Vym, s->Np, s->Nm, s->Op, s->Om, s->Vx, s->Vy,
        s->Zx, s->Zy, s->Nx, s->Oy, s->minVxp, s->minVxm, s->minVyp, s->minVym, s->minVx,
        s->minVy, s->Nmin, s->Omin, s->Nmax, s->Omax, s->tYhX, s->tYhY, s->tCn, s->YhX, s->YhY, s->Cn,
        s->tNmin, s->tNmax, s->tOmin, s->tOmax};
    for (size_t q = 0; q < sizeof(subs) / sizeof(subs[0]); q++) free(subs[q]);
    free(s->ind);
    free(s->tideLocP); free(s->tideLocM); free(s->inflowLoc);
    free(s);
}

double *fo_field(fo_state *s, const char *name, int64_t *len)
{
    struct { const char *n; double *p; int64_t l; } tab[] = {
        {"surf", s->surf, s->Ng}, {"surfOld", s->surfOld, s->Ng}, {"uuXP", s->uuXP, s->Ng},
        {"vvYP", s->vvYP, s->Ng}, {"uuYP", s->uuYP, s->N}, {"vvXP", s->vvXP, s->N},
        {"depth", s->depth, s->N}, {"depthXP", s->depthXP, s->Ng}, {"depthYP", s->depthYP, s->Ng},
        {"EnXP", s->EnXP, s->Ng}, {"EnYP", s->EnYP, s->Ng}, {"DragXP", s->DragXP, s->Ng},
        {"DragYP", s->DragYP, s->Ng}, {"CDXP", s->CDXP, s->Ng}, {"CDYP", s->CDYP, s->Ng},
        {"Fuu", s->Fuu, s->Ng}, {"Fvv", s->Fvv, s->Ng}, {"cellV", s->cellV, s->N},
        {"S", s->S, (int64_t)s->cfg.nScalar * s->Ng}, {"Sm", s->Sm, (int64_t)s->cfg.nScalar * s->Ng},
        {"CFLx", s->CFLx, s->Ng}, {"CFLy", s->CFLy, s->Ng}, {"wtfXP", s->wtfXP, s->Ng},
        {"wtfYP", s->wtfYP, s->Ng}, {"bottomZ", s->bottomZ, s->Ng}, {"bottomXP", s->bottomXP, s->Ng},
        {"bottomYP", s->bottomYP, s->Ng}, {"advX", s->advX, s->Ng}, {"advY", s->advY, s->Ng},
        {"rmvCFL", s->rmvCFL, s->N}, {"GnCt", s->GnCt, s->N}, {"GnXP", s->GnXP, s->N},
        {"GnXM", s->GnXM, s->N}, {"GnYP", s->GnYP, s->N}, {"GnYM", s->GnYM, s->N},
        {"z", s->z, s->N}, {"V", s->V, s->Ng}, {"Z", s->Z, s->Ng}, {"Vx", s->Vx, s->Ng},
        {"Vy", s->Vy, s->Ng}, {"Zx", s->Zx, s->Ng}, {"Zy", s->Zy, s->Ng}, {"Nx", s->Nx, s->Ng},
        {"Oy", s->Oy, s->Ng}, {"Vxp", s->Vxp, s->N}, {"Vxm", s->Vxm, s->N}, {"Vyp", s->Vyp, s->N},
        {"Vym", s->Vym, s->N}, {"Np", s->Np, s->N}, {"Nm", s->Nm, s->N}, {"Op", s->Op, s->N},
        {"Om", s->Om, s->N}, {"minVx", s->minVx, s->Ng}, {"minVy", s->minVy, s->Ng}};
    for (size_t q = 0; q < sizeof(tab) / sizeof(tab[0]); q++)
        if (strcmp(tab[q].n, name) == 0 && tab[q].p) { if (len) *len = tab[q].l; return tab[q].p; }
    if (len) *len = 0;
    return NULL;
}

int32_t *fo_ind(fo_state *s) { return s->ind; }

int32_t fo_tt(const fo_state *s) { return s->tt; }
int32_t fo_pcg_iters(const fo_state *s) { return s->pcgIters; }
int64_t fo_pcg_iters_total(const fo_state *s) { return s->pcgItersTotal; }
double  fo_pcg_relres(const fo_state *s) { return s->relres; }
void    fo_vloss(const fo_state *s, double out[4]) { for (int q = 0; q < 4; q++) out[q] = s->Vloss[q]; }
