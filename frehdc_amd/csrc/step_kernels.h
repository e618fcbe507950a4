// step_kernels.h -- argument block shared by the step kernels and the host
// orchestration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frehd_dev.h"

namespace frehd {

// scalar settings (the Config fields the step reads)
struct P {
    double dx, dy, dt, g, NUx, NUy, CDx, CDy, manningN, hD, CDmax, minDepth, wtfh, CFLl, CFLh;
    double eps, tidePS, tideMS, inflowS, Kx, Ky, offset;
    int CDnotN, useThinLayer, bcType, tideLocLengthP, tideLocLengthM, inflowLocLength;
    int useScalar, nScalar, useConstTidePS, useConstTideMS, maxIter, useSubgrid;
    // subgrid
    double dsurf, subR;
    int Ne;
    int warmStart;            // mixed precision: refinement starts from eta^n instead of 0
};

struct PcgScalars {
    double bb, bNorm, rr, rNorm, pq, alpha;
    double rhoHist[2];        // r.z of iterations of even / odd index
    int iter;
    int doneAt;               // iteration whose stop test succeeded, -1 while running
};

// FP64 refinement of the mixed-precision solve
struct RefineScalars {
    double bNorm, rNorm;
    double epsIn;             // relative tolerance of the current inner FP32 solve
    int sweep;                // last refinement sweep tested
    int doneAt;               // sweep whose residual test succeeded, -1 while running
    int innerIters[8];        // inner iterations of sweeps 0..7 of the last solve
};

// device control block (one per handle)
struct Ctl {
    int tt;                   // current step number
    int halted;               // set when max CFL > 20
    int haltStep;
    int pad;
    double lastMaxCFLx, lastMaxCFLy;
    double Vloss[4];
    PcgScalars pcg;
    int pcgIterLast;
    int pad0;
    long long pcgIterTotal;
    long long steps;
    double relResLast;
    RefineScalars ref;
    PcgScalars pcgIn;         // inner FP32 solve of the mixed-precision path
    int refineLast;
    int pad1;
};

struct StepArgs {
    Grid g;
    P p;
    Ctl *ctl;
    // Ng arrays
    double *surf, *surfOld, *uuXP, *vvYP, *depthXP, *depthYP, *EnXP, *EnYP, *DragXP, *DragYP;
    double *CDXP, *CDYP, *Fuu, *Fvv, *CFLx, *CFLy, *wtfXP, *wtfYP, *bottomZ, *bottomXP, *bottomYP;
    double *S, *Sm;           // nScalar arrays with stride sStride
    int64_t sStride;
    // interior arrays
    double *uuYP, *vvXP, *depthOld, *depthNew, *cellV, *Snew, *manning;
    uint8_t *rmv, *flags;     // flags: bit0 tideP cell, bit1 tideM cell, bits 2-7 inflow count
    // operator and solver vectors
    double *cX, *cY, *D, *b, *x, *r, *q, *pA, *pOld, *pNew;
    float *cX32, *cY32, *D32;  // FP32 operator copies (mixed precision only, else NULL)
    // per-block partial sums: solver (partA: r.r, r.z; partB: p.q), volume
    // losses (vpart: 4 slots) and CFL maxima (cflpart: 2 slots), each slot
    // maxBlocks long; nb1 = blocks of an Ng sweep, nb2 = blocks of a 2-D sweep
    double *partA, *partB, *vpart, *cflpart;
    int maxBlocks, nb1, nb2, pad3;
    // forcing (Nt+1 samples) and locations
    const double *tideP, *tideM, *inflow, *tidalPS, *tidalMS;
    const int *tideLocP, *tideLocM;
    // subgrid: state (Ng), half-cell values (N), level search index (Ng),
    // face bottoms with offset (N), level elevations (Ne), level tables
    // V Z Vxp Vxm Vyp Vym Np Nm Op Om interleaved per cell and level:
    // tabs[(c*Ne + k)*10 + T], so the two levels a cell interpolates between
    // are 160 contiguous bytes (the reference's files are level-major,
    // [k*N + c], and are reordered on upload)
    double *V, *Z, *Vx, *Vy, *Zx, *Zy, *Nx, *Oy;
    double *Vxp, *Vxm, *Vyp, *Vym, *Np, *Nm, *Op, *Om;
    int *ind;
    const double *edgeX, *edgeY, *allSurf;
    const double *tabs;
};

__host__ __device__ inline size_t tab_index(size_t c, int k, int Ne, int T)
{
    return (c * (size_t)Ne + (size_t)k) * 10 + (size_t)T;
}

// One Jacobi-PCG solve over device vectors of type T (double for the FP64
// path, float for the inner solve of the mixed-precision path).  Operator
// row c (= j*nx + i): W -cX[c-1], S -cY[c-nx], C D[c], N -cY[c], E -cX[c];
// D < 0 marks an identity row scaled by -D (tidal Dirichlet rows, isolated
// dry subgrid cells).
template <typename T>
struct PcgArgs {
    const T *cX, *cY, *D;
    T *r, *q, *pOld, *pNew, *x;
    double *partA;            // r.r, r.z partials: of the start residual (nbStart
                              // blocks) at iteration 0, of k_pcg_res (nbRes) after
    double *partB;            // p.q partials of k_pcg_dir (nbDir blocks)
    PcgScalars *sc;
    const int *halted;
    const int *outerDone;     // inner solve: the refinement has converged (>= 0)
    const double *epsDyn;     // inner solve: tolerance set by the refinement test
    double eps;
    int maxIter, nx, ny, N;
    // k_pcg_dir tiling: gx x gy tiles of 256 columns x H rows; remap != 0
    // gives each XCD a contiguous band of tiles (nbDir a multiple of 8)
    int gx, gy, H, remap;
    int nbDir, nbRes, nbStart;
    int decouple;             // inner solve: skip couplings to identity rows
};

template <typename T> __global__ void k_pcg_dir(PcgArgs<T> a, int it);
template <typename T> __global__ void k_pcg_res(PcgArgs<T> a, int it);

// FP64 iterative refinement around the FP32 solve (mixed precision): sweep
// s forms x_s = x_{s-1} + d (d: the inner solve's correction) and the true
// residual r = b - A x_s with the FP64 operator; k_refine_test decides
// whether the solve has converged and sets up the next inner solve.
struct RefineArgs {
    const double *cX, *cY, *D, *b;
    double *xbuf[2];          // x_s lives in xbuf[s & 1]
    const float *d;           // inner correction
    float *r32;               // residual, rounded: the next inner right-hand side
    const float *D32;
    double *partA;            // r32.r32, r32.z32 partials (inner start)
    double *partC;            // b.b, r.r partials (FP64)
    Ctl *ctl;
    double eps, epsInner;
    int maxRefine, nx, ny, N;
    int gx, gy, H, remap, nb, pad;
};

__global__ void k_refine(RefineArgs a, int s);
__global__ void k_refine_test(RefineArgs a, int s);

enum SubTable { T_V, T_Z, T_VXP, T_VXM, T_VYP, T_VYM, T_NP, T_NM, T_OP, T_OM };

__global__ void k_step_begin(StepArgs a);
__global__ void k_predictor(StepArgs a);
__global__ void k_assemble(StepArgs a);
__global__ void k_surface(StepArgs a);
__global__ void k_limiter(StepArgs a);
__global__ void k_faces(StepArgs a);
__global__ void k_flux(StepArgs a);
__global__ void k_tidal_velocity(StepArgs a);
__global__ void k_finish(StepArgs a);
__global__ void k_scalar(StepArgs a, int q);
__global__ void k_scalar_bc(StepArgs a, int q);
__global__ void k_step_end(StepArgs a);
// subgrid variants
__global__ void k_predictor_sub(StepArgs a);
__global__ void k_assemble_sub(StepArgs a);
__global__ void k_subarea_init(StepArgs a);
__global__ void k_limiter_sub(StepArgs a);
__global__ void k_combine(StepArgs a);
__global__ void k_sub_ghosts(StepArgs a);
__global__ void k_faces_sub(StepArgs a);
__global__ void k_flux_sub(StepArgs a);
__global__ void k_tidal_velocity_sub(StepArgs a);
__global__ void k_finish_sub(StepArgs a);
__global__ void k_scalar_sub(StepArgs a, int q);

}  // namespace frehd
