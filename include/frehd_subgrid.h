/*
 * frehd_subgrid.h -- the subgrid level tables of FrehdC's pre-processing,
 * generated from a fine bathymetry on the host, and written as the text
 * files the reference's initAllSubVar reads (FrehdC_v42/subgrid.c:2088-2284,
 * ReadFile fileio.c:28-42: one value per line).
 *
 * Restates preProcessing/script_CreateSubAreaWithBlockV16.m:144-460 with
 * ComputeFaceAreaWithBlockCheck.m:1-326 (checkBlock = 1) or with the plain
 * face areas of ComputeSubgridFaceArea.m (checkBlock = 0), and the writer of
 * preProcessing/script_GenerateInput.m:174-220 ("%f\n" per value, the
 * script's column-major order, which is the reference's level-major order
 * t[k*N + j*nx + i]).
 *
 * Orientation: the script's first index is the reference's x index i (its
 * arrays are reshaped column-major); here the fine bathymetry is row-major
 * fine[Y * nxFine + X] with X along x, and "p"/"m" halves and faces are the
 * +x (N) and +y (O) sides, as in the tables the step reads.
 *
 * Levels are surfmin + k * dsurf, k = 0 .. Ne-1 (the script's colon range
 * holds Ne + 1 values; the reference reads Ne = round((surfmax - surfmin) /
 * dsurf) of them, subgrid.c:2116).
 */
#ifndef FREHD_SUBGRID_H
#define FREHD_SUBGRID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct frehd_subgrid_spec {
    int32_t nxFine, nyFine;   /* pixels; nxFine = r * nx, nyFine = r * ny, nx, ny >= 2      */
    int32_t r;                /* pixels per coarse cell side (even with checkBlock = 1)      */
    int32_t Ne;               /* levels                                                      */
    double  dxf;              /* pixel side (the script's dx)                                */
    double  surfmin, dsurf;
    int32_t checkBlock;       /* 1: face blocking, minimum / maximum face areas              */
    int32_t useminVolume;     /* 1: minimum half-cell volumes (the script's useminVolume)    */
    int32_t useYh;            /* 1: YhX, YhY of the Volp drag model (useYh)                  */
    int32_t useCn;            /* 1: Cn of the Casas model (useCn)                            */
    double  n0;               /* Cn outside the model's range (the script's n0)              */
} frehd_subgrid_spec;

/* Outputs, each Ne * nx * ny doubles (level-major) except the three N-sized
 * geometry arrays; a NULL pointer skips that output. */
typedef struct frehd_subgrid_out {
    double *V, *Z, *Vxp, *Vxm, *Vyp, *Vym, *Np, *Nm, *Op, *Om;
    double *Nmin, *Nmax, *Omin, *Omax;            /* checkBlock = 1 (zeros otherwise)  */
    double *minVxp, *minVxm, *minVyp, *minVym;    /* useminVolume = 1                  */
    double *YhX, *YhY, *Cn;                       /* useYh / useCn = 1                 */
    double *bottom, *bottomXP, *bottomYP;         /* cell and face bottoms (N each)    */
} frehd_subgrid_out;

/* Generate the tables; levels are spread over `threads` host threads (0: as
 * many as the machine has).  FREHD_EINVAL for inconsistent sizes. */
int frehd_subgrid_generate(const double *fine, const frehd_subgrid_spec *spec, frehd_subgrid_out *out,
                           int32_t threads);

/* Write every non-NULL output as "<folder>/<name>.dat", one "%f" line per
 * value (names: V Z Vxp Vxm Vyp Vym Np Nm Op Om Nmin Nmax Omin Omax minVxp
 * minVxm minVyp minVym YhX YhY Cn bottom bottomXP bottomYP); the folder is
 * created when missing. */
int frehd_subgrid_write_text(const char *folder, const frehd_subgrid_spec *spec, const frehd_subgrid_out *t);

#ifdef __cplusplus
}
#endif

#endif /* FREHD_SUBGRID_H */
