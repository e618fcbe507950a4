/*
 * frehd_io.h -- FrehdC v4.2 input decks and output files for the device step.
 *
 * Reads what the reference's main program reads before its time loop
 * (FrehdC_v42/FREHD.c:28-60 ReadInputFile, configuration.c:18-158
 * ReadUserSettings, bathymetry.c:19-104 ReadBathymetry/InitBathymetry,
 * initialize.c:280-460 initial conditions and readBC/readOneBC with
 * utilities.c:51-85 dataSplit/dataInterp and :129-150 searchInd,
 * subgrid.c:2085-2284 initAllSubVar) and writes what its DataOutput and
 * SubOutput write (fileio.c:62-180, WriteFile :181-206): plain-text files
 * "<output folder><name>_<step>.dat", one "%4.4f \n" line per interior
 * cell.  The loaded case hands the device library its frehd_config,
 * frehd_forcing and frehd_ic directly.
 *
 * Differences from the reference, each a defined outcome where the
 * reference's is undefined or a crash:
 *   - the start date "YYYY-MM-DD" is parsed completely (the reference copies
 *     the digits into an unterminated buffer, utilities.c:114-123) and
 *     converted with mktime in the process's local time zone, as the
 *     reference does (utilities.c:98-110);
 *   - boundary series get Nt + 1 samples, the last one from the same
 *     interpolation (the reference allocates Nt and reads index Nt,
 *     initialize.c:333-341 vs nsfunctions.c:386);
 *   - a time index the series does not contain fails with FREHD_EINVAL
 *     (the reference prints a warning and reads past the array,
 *     utilities.c:137-148);
 *   - the output folder is created when missing (the reference's fopen
 *     fails there).
 */
#ifndef FREHD_IO_H
#define FREHD_IO_H

#include <stdint.h>

#include "frehd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct frehd_case frehd_case_t;

/* what the time loop needs besides the step: output cadence and switches */
typedef struct frehd_case_info {
    int32_t Nt;              /* steps: NT days * 86400 / dt, truncated            */
    int32_t OutItvl;         /* output every OutItvl steps (hours * 3600 / dt)    */
    int32_t savesurface, saveuu, savevv, savedepth, savescalar, saveCD, savesub;
    int32_t useSubgrid;
    double  offset;          /* Bath.offset[0]                                    */
    double  tNStart;         /* start date as seconds since the epoch             */
    char    outputFolder[512];
} frehd_case_info;

/* Load the deck `data_dat`; folders named in it are relative to base_dir
 * (the reference's working directory).  The case owns every array it hands
 * out; it must outlive the handles created from it. */
int frehd_case_load(const char *data_dat, const char *base_dir, frehd_case_t **out);

/* The device inputs of the case (pointers into the case). */
int frehd_case_inputs(const frehd_case_t *c, frehd_config *cfg, frehd_forcing *forcing, frehd_ic *ic);

int frehd_case_get_info(const frehd_case_t *c, frehd_case_info *info);

/* DataOutput (and, with useSubgrid, SubOutput) of step tt from a handle:
 * the arrays the save switches select, surface minus the offset. */
int frehd_case_write_output(const frehd_case_t *c, frehd_dev_t *dev, int32_t tt);

/* The same from host arrays (Ng-layout surf/uuXP/vvYP/CDXP/CDYP, N-layout
 * depth, scalar 0 in S; subgrid Nx Oy Vx Vy V Z, Ng each); NULL skips. */
int frehd_case_write_arrays(const frehd_case_t *c, int32_t tt, const double *surf, const double *uuXP,
                            const double *vvYP, const double *depth, const double *CDXP,
                            const double *CDYP, const double *S, const double *Nx, const double *Oy,
                            const double *Vx, const double *Vy, const double *V, const double *Z);

int frehd_case_free(frehd_case_t *c);

#ifdef __cplusplus
}
#endif

#endif /* FREHD_IO_H */
