/*
 * frehd_oracle.h -- CPU restatement of FrehdC v4.2's semi-implicit step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the device path is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product library never links or calls it.
 *
 * It follows the reference's sequential order of floating-point operations so
 * that, on the same inputs, it reproduces the reference bit for bit (checked
 * against the reference's own code built by oracle/Makefile into oracle/_ref).
 * Parity status of each part is stated in oracle/frehd_oracle.c.
 */
#ifndef FREHD_ORACLE_H
#define FREHD_ORACLE_H

#include "../include/frehd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fo_state fo_state;

/* phases of one step, in the order oneCompleteStep runs them (nssolve.c:110-118) */
#define FO_PHASE_SOURCE   0   /* solveSourceTerm   nssolve.c:121-155 */
#define FO_PHASE_SURFACE  1   /* solveFreeSurface  nssolve.c:158-178 */
#define FO_PHASE_UPDATE   2   /* updateData        nssolve.c:181-290 */
#define FO_PHASE_SCALAR   3   /* scalarTransport   scalar.c:30-56     */
#define FO_PHASE_CFL      4   /* monitorCFL        nsfunctions.c:1192 */

fo_state *fo_create(const frehd_config *cfg, const frehd_forcing *forcing, const frehd_ic *ic);
void      fo_destroy(fo_state *s);
/* one full step (all phases); returns 0 or FREHD_ECFL */
int       fo_step(fo_state *s);
/* run one phase of step s->tt (FO_PHASE_SOURCE increments tt first) */
int       fo_phase(fo_state *s, int phase);
/* pointer to a named array and its length in doubles (NULL if unknown);
   names are the frehd_fields member names plus "z" (RHS, cell order),
   "GnCt", "GnXP", "GnXM", "GnYP", "GnYM", "advX", "advY", "rmvCFL" */
double   *fo_field(fo_state *s, const char *name, int64_t *len);
int32_t   fo_tt(const fo_state *s);
int32_t   fo_pcg_iters(const fo_state *s);
double    fo_pcg_relres(const fo_state *s);
void      fo_vloss(const fo_state *s, double out[4]);
int64_t   fo_pcg_iters_total(const fo_state *s);

/* The free-surface solve alone, for operator tests: CG on the 5-point
 * system with the given row coefficients in cell order (the matrix that
 * setupMatrix builds, nsfunctions.c:927-1001), solved in the reference's
 * transposed row order.  Returns the iteration count. */
int       fo_solve(int nx, int ny, const double *GnCt, const double *GnXP, const double *GnXM,
                   const double *GnYP, const double *GnYM, const double *z, double *x,
                   double eps, int maxIter, int precond, double *relres);

#ifdef __cplusplus
}
#endif

#endif
