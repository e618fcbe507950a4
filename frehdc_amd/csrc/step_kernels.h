// step_kernels.h -- argument block shared by the step kernels and the host
// orchestration.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frehd_dev.h"
#include "subgrid_eval.h"

namespace frehd {

// scalar settings (the Config fields the step reads)
struct P {
    double dx, dy, dt, g, NUx, NUy, CDx, CDy, manningN, hD, CDmax, minDepth, wtfh, CFLl, CFLh;
    double eps, tidePS, tideMS, inflowS, Kx, Ky, offset;
    int CDnotN, useThinLayer, bcType, tideLocLengthP, tideLocLengthM, inflowLocLength;
    int useScalar, nScalar, useConstTidePS, useConstTideMS, maxIter, useSubgrid;
    int scalarAdv;            // 1 first-order upwind, 2 TVD superbee
    int useWind;
    double rhoa, Cw, CwT;     // wind stress (windTerm)
    int useSubDrag;           // 0 Manning h = V/Z, 2 Volp (Yh), 3/4 Casas (Cn)
    int useminA, useminV;     // minimum face areas / effective volumes
    int minTabs;              // the minimum-area and minV tables are on the device
    double beta;
    // subgrid
    double dsurf, subR, surfmin;   // surfmin: raw elevation of level 0
    int Ne;
    int warmStart;            // mixed precision: refinement starts from eta^n instead of 0
    int budget;               // account the conservation budgets (frehd_dev_set_budget)
};

// Flux-form conservation budgets, accumulated over the steps since they
// were switched on: per step, the change of the domain's volume (of each
// scalar's mass) minus what crossed the domain boundary, the inflow and --
// for the volume -- what the clip of negative cell volumes added; sums in
// double-double per step (the residual), then in FP64 over steps.
constexpr int kMaxScalars = 16;
struct Budget {
    long long steps;
    double vol, volRes, volResMax, volBnd, volIn, volClip, volMoved;
    double mass[kMaxScalars], massRes[kMaxScalars], massResMax[kMaxScalars];
    double massBnd[kMaxScalars], massIn[kMaxScalars], massMoved[kMaxScalars];
};
// per-block partial slots of the budgets (bpart[slot * maxBlocks + block]):
// volumeByFlux's 7 (residual hi, lo, volume, boundary, inflow, clip, sum of
// the cells' |change|), then 6 per scalar (residual hi, lo, mass, boundary,
// inflow, sum of |change|)
constexpr int kBudgetVolSlots = 7, kBudgetScalarSlots = 6;
__host__ __device__ inline int budget_slots(int nScalar) { return kBudgetVolSlots + kBudgetScalarSlots * nScalar; }

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
    int innerConv;            // bit s: the inner solve after sweep s met its own stop test
};

// device control block (one per handle)
// (kernels read their member's argument block from a device array,
// A[blockIdx.z]: one launch can advance many scenarios)
struct Ctl {
    int tt;                   // current step number
    int halted;               // set when max CFL > 20
    int haltStep;
    int parity;               // depth buffer of the step's start: 0 -> depthA
    int sParity;              // scalar buffer of the step's start: 0 -> S
    int open;                 // the step's solve has not finished (set by the flag gather)
    int runTail;              // the kernels after the solve run for this member
                              // (set by the gather that first sees its solve done)
    int tailIssued;           // that gather has run in this step (cleared at the step's begin)
    int flipPending;          // the update wrote the new depth (and the scalar transport the
                              // concentrations) into the second buffers: the next step's
                              // begin flips parity / sParity
    int haltPending;          // a cell's CFL exceeded 20 in this step (set by the velocity
                              // update); the next step's begin halts the member
    int ttNext;               // tt + 1 at every step boundary: the step the next
                              // predictor computes (its first thread advances tt
                              // while the other blocks run, so they read this)
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
    Budget budget;
};

// The explicit phase's state arrays are stored in R: FP64 for the FP64 form,
// FP32 for the mixed form (which keeps eta, the bed, the cell depths and
// volumes, the scalar masses, the solver vectors and the subgrid lookup's
// surface cache in FP64).  Arithmetic is FP64 in both (sptr): an FP32 value
// read is widened, a result stored is rounded to R.  The host
// fills StepArgs (R = double): the pointers of both instantiations have
// the same size and place, so the device reads the same block as
// StepArgsT<float> when the handle stores FP32.
template <class R>
struct StepArgsT {
    Grid g;
    P p;
    gptr<Ctl> ctl;
    // Ng arrays; the free surface is double-buffered (sf_start / sf_new:
    // Ctl.parity selects, as for the depth)
    gptr<double> surfA, surfB;
    sptr<R> uuXP, vvYP, depthXP, depthYP, EnXP, EnYP, DragXP, DragYP;
    sptr<R> CDXP, CDYP, Fuu, Fvv, CFLx, CFLy, wtfXP, wtfYP;
    gptr<double> bottomZ, bottomXP, bottomYP;
    // nScalar arrays with stride sStride; the concentration is
    // double-buffered (S, S2; Ctl.sParity selects the step's start), the
    // scalar mass Sm is not
    sptr<R> S, S2;
    gptr<double> Sm;
    int64_t sStride;
    // interior arrays; the depth is double-buffered: Ctl.parity selects the
    // buffer holding the depth at the start of the step
    sptr<R> uuYP, vvXP;
    gptr<double> depthA, depthB, cellV, manning;
    gptr<uint8_t> rmv, flags;      // flags: bit0 tideP cell, bit1 tideM cell, bits 2-7 inflow count
    gptr<uint16_t> tface;          // per cell: which of the vvYP faces its velocity interpolation
                                   // reads the tidal adjustment sets, and how (bit pairs 0-7),
                                   // bit 8 its jM ghost face
    // operator and solver vectors
    gptr<double> cX, cY, D, b, x, r, q, pA, pB;
    gptr<float> cX32, cY32;        // mixed precision only (else NULL): the scaled FP32
                                   // operator of the inner solve, written by k_refine
    gptr<double> xB;               // mixed precision: second refinement x buffer
    // per-block partial sums: solver (partA: r.r, r.z; partB: p.q), volume
    // losses (vpart: 4 slots) and CFL maxima (cflpart: 2 slots), each slot
    // maxBlocks long; nb1 = blocks of an Ng sweep, nb2 = blocks of a 2-D sweep
    gptr<double> partA, partB, vpart, cflpart;
    gptr<double> bpart;            // conservation-budget partials (budget_slots x maxBlocks)
    int maxBlocks, nb1, nb2, pad3;
    // this member's launch extents (a batched launch is sized for its
    // largest member; blocks past these exit at once): 2-D stencil grid of
    // 64x4 blocks, sweeps of 256 over N, ghost cells and tidal cells, and
    // the grid-stride solver grid
    int gx2, gy2, nbN, nbGhost, nbTide, nbRes;
    int rowsDown;                  // the 2-D kernels' blocks take rows from the last down (row_block)
    int nbAsm;                     // the assembly's blocks (nbRes when it writes the solver's partials)
    // forcing (Nt+1 samples) and locations
    gptr<const double> tideP, tideM, inflow, tidalPS, tidalMS;
    // wind per step: speed and the cosine / sine of its direction (computed
    // on the host with the reference's expression and libm)
    gptr<const double> windSpd, windCos, windSin;
    gptr<const int> tideLocP, tideLocM;
    // subgrid: state (Ng), half-cell values (N), level search index (Ng),
    // face bottoms with offset (N), level elevations (Ne), level tables
    // V Z Vxp Vxm Vyp Vym Np Nm Op Om interleaved per cell and level:
    // tabs[(c*Ne + k)*10 + T], so the two levels a cell interpolates between
    // are 160 contiguous bytes (the reference's files are level-major,
    // [k*N + c], and are reordered on upload)
    sptr<R> V, Z, Vx, Vy, Zx, Zy, Nx, Oy;
    sptr<R> Vxp, Vxm, Vyp, Vym, Np, Nm, Op, Om;
    // computeSubArea's own V and Z (before combineSubArea clips them in
    // place), and the inputs it last saw per cell (surface, dry-start flag):
    // a cell whose inputs are bitwise unchanged keeps its outputs
    sptr<R> Vraw, Zraw;
    gptr<double> hSub;
    gptr<uint8_t> dSub;
    gptr<int> ind;
    gptr<const double> edgeX, edgeY, allSurf;
    gptr<const R> tabs;            // in the state precision (FP32 for the mixed form)
    // drag-model tables (level-major [k*N + c]) and the values at the lower
    // bracketing level of each cell (useSubDrag 2: Yh; 3, 4: Cn)
    gptr<const double> allYhX, allYhY, allCn;
    gptr<double> YhX, YhY, Cn;
    // minimum / maximum face areas and minimum half-cell volumes (level-major
    // tables, per-cell values of the step) when the combine uses them
    gptr<const double> allNmin, allNmax, allOmin, allOmax, allMinVxp, allMinVxm, allMinVyp, allMinVym;
    gptr<double> Nmin, Nmax, Omin, Omax, minVxp, minVxm, minVyp, minVym;
    // instead of tabs (NULL): the fine bathymetry, from which a lookup
    // evaluates its two levels (cell_levels, bitwise the table values)
    FineBed bed;
};
using StepArgs = StepArgsT<double>;


// value T of cell c at level k: levels outermost, the ten values of a cell
// adjacent, cells in order within a level.  A wave's cells sit at about the
// same level (the water surface is smooth), so a lookup's two levels are
// two contiguous runs of 80 bytes per cell.
__host__ __device__ inline size_t tab_index(size_t c, int k, size_t N, int T)
{
    return ((size_t)k * N + c) * 10 + (size_t)T;
}

// Loads and stores through the global address space from a plain pointer
// (the same reason as gptr).
#define FREHD_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ inline T gld(const T *p)
{
    return *(const FREHD_GLOBAL T *)p;
}
template <typename T>
__device__ inline void gst(T *p, const T &v)
{
    *(FREHD_GLOBAL T *)p = v;
}

// One Jacobi-PCG solve over device vectors of type T (double for the FP64
// path, float for the inner solve of the mixed-precision path).  Operator
// row c (= j*nx + i): W -cX[c-1], S -cY[c-nx], C D[c], N -cY[c], E -cX[c];
// D < 0 marks an identity row scaled by -D (tidal Dirichlet rows, isolated
// dry subgrid cells).
template <typename T>
struct PcgArgs {
    gptr<const T> cX, cY, D;
    gptr<T> r, q, x;
    gptr<T> pBuf[2];               // iteration it reads p from pBuf[it&1], writes pBuf[(it+1)&1]
    gptr<double> partA;            // r.r, r.z (, b.b) partials: of the start residual
                                   // (nbStart blocks) at iteration 0, of k_pcg_res (nbRes) after
    gptr<double> partB;            // p.q partials of k_pcg_dir (nbDir blocks)
    gptr<PcgScalars> sc;
    gptr<const int> halted;
    gptr<const int> outerDone;     // inner solve: the refinement has converged (>= 0)
    gptr<const double> epsDyn;     // inner solve: tolerance set by the refinement test
    double eps;
    int maxIter, nx, ny, N;
    // k_pcg_dir / k_cg1 tiling: gx x gy tiles of 256 lanes columns x H rows
    // (lanes: adjacent columns per lane, k_cg1 only; 1 for k_pcg_dir); remap
    // != 0 gives each XCD a contiguous band of tiles (nbDir a multiple of 8)
    int gx, gy, H, remap, lanes;
    int nbDir, nbRes, nbStart;
    gptr<double> partF[2];         // k_cg1: p.q, r.q, q.q, r.r partials (nbDir each);
                                   // k_pcg1: p.q, z.q, q.M^-1 q, r.r, r.z; launch it
                                   // writes partF[it&1]
    // SSOR preconditioning (FP64 only): k_ssor writes z = M^-1 r here and
    // r.z into sc->rhoHist[it & 1]; k_pcg_dir then reads z instead of
    // forming r / |D|
    gptr<T> z;
    int ssor, ldsCells;            // ldsCells: k_ssor keeps z in LDS when N <= this
    // FP64 Jacobi from a nonzero start x_0 (warm != 0): k_pcg1_start forms
    // r = b - A x_0 and the start partials
    gptr<const T> b;
    int warm, pad;
};

// The tile a solver block walks: with remap, launch block vb takes tile
// (vb mod 8) * nb/8 + vb/8 (nb a multiple of 8), so the blocks the
// dispatcher hands to one XCD walk one contiguous band of tiles; block vb
// writes its partials into slot vb.  slot_of is the inverse: the slot of
// tile lb's partials.
__host__ __device__ inline int tile_of(int vb, int nb, int remap)
{
    return remap ? (vb & 7) * (nb >> 3) + (vb >> 3) : vb;
}
__host__ __device__ inline int slot_of(int lb, int nb, int remap)
{
    return remap ? (lb % (nb >> 3)) * 8 + lb / (nb >> 3) : lb;
}

template <typename T> __global__ void k_pcg_dir(const PcgArgs<T> *__restrict__ A, int it);
template <typename T> __global__ void k_pcg_res(const PcgArgs<T> *__restrict__ A, int it);
// SSORPrecond with omega = 1 for iteration it, one workgroup per member
// (LDS: the sweep's vector in dynamic LDS, else in z itself)
template <bool LDS> __global__ void k_ssor(const PcgArgs<double> *__restrict__ A, int it);
struct RefineArgs;
// the inner FP32 solve of the mixed path, one launch per iteration (First:
// the set-up launch, it = 0; W = lanes columns per lane): cX, cY hold the
// scaled coefficients, r and q the two residual buffers, D unused.  The
// set-up launch first runs the outer test of refinement sweep testS when
// testS >= 0 (R: the members' refinement arguments)
template <bool First, int W>
__global__ void k_cg1(const PcgArgs<float> *__restrict__ A, int it, const RefineArgs *__restrict__ R, int testS);
// the FP64 Jacobi PCG, one launch per iteration (First: launch 0; W = lanes
// columns per lane, 1 or 2): partials in partF, r in r / q and p in pBuf
// alternating by iteration
template <bool First, int W> __global__ void k_pcg1(const PcgArgs<double> *__restrict__ A, int it);
// r = b - A x_0 and the start partials (r.r, r.z, b.b) of a member with warm != 0
__global__ void k_pcg1_start(const PcgArgs<double> *__restrict__ A);

// FP64 iterative refinement around the FP32 solve (mixed precision): sweep
// s forms x_s = x_{s-1} + d (d: the inner solve's correction) and the true
// residual r = b - A x_s with the FP64 operator; refine_test (run by the
// launch after it) decides whether the solve has converged and sets up the
// next inner solve.
struct RefineArgs {
    gptr<const double> cX, cY, D, b;
    gptr<double> xbuf[2];          // x_s lives in xbuf[s & 1]
    gptr<const float> d;           // inner correction
    gptr<float> r32;               // scaled residual, rounded: the next inner right-hand side
    gptr<float> cXs, cYs;          // scaled FP32 coefficients of the inner solve (written at s = 0)
    gptr<double> partA;            // r32.r32 partials (inner start)
    gptr<double> partC;            // b.b, r.r partials (FP64)
    gptr<Ctl> ctl;
    double eps, epsInner;
    int maxRefine, nx, ny, N;
    int gx, gy, H, remap, nb, pad;
};

__global__ void k_refine(const RefineArgs *__restrict__ A, int s, const PcgArgs<float> *__restrict__ P);

// The outer stop test after refinement sweep s: |b - A x_s| < eps |b| (the
// reference's criterion, rtc.c:54, on the FP64 residual) or s at the sweep
// cap; otherwise the next inner solve's tolerance max(epsInner,
// eps |b| / (2 |r|)), so the last inner solve stops about when the outer
// test will pass.  Every thread of the block takes part (the partials of
// k_refine(s) are reduced in a fixed order, so every block that runs it
// obtains the same result); thread 0 of the block with `record` writes it
// to the control block.  Run by the launch after k_refine(s): the inner
// solve's set-up launch, or the flag gather.
struct RefTest {
    bool conv;
    double epsIn;
};

__device__ inline RefTest refine_test(const RefineArgs &a, int s, double *lds /* 16*2 */, bool record)
{
    Ctl *ctl = a.ctl;
    double v[2];
    reduce_partials<2>(a.partC, a.nb, v, lds);
    RefineScalars &R = ctl->ref;
    const double bNorm = s == 0 ? sqrt(v[0]) : R.bNorm, rNorm = sqrt(v[1]);
    const bool conv = rNorm < a.eps * bNorm || (LAS_IS_ZERO(bNorm) && LAS_IS_ONE(1.0 + rNorm)) ||
                      s >= a.maxRefine;
#ifndef FREHD_INNER_MARGIN
#define FREHD_INNER_MARGIN 0.5
#endif
    const double epsIn = fmax(a.epsInner, FREHD_INNER_MARGIN * a.eps * bNorm / rNorm);
    if (record && threadIdx.x == 0) {
        if (s == 0) R.bNorm = bNorm;
        if (s > 0) {
            const int n = ctl->pcgIn.iter;
            if (s - 1 < 8) R.innerIters[s - 1] = n;
            if (s - 1 < 31 && ctl->pcgIn.doneAt >= 0) R.innerConv |= 1 << (s - 1);
            ctl->pcg.iter += n;
        }
        R.rNorm = rNorm;
        R.sweep = s;
        if (conv) {
            R.doneAt = s;
            ctl->pcg.rNorm = rNorm;
            ctl->pcg.bNorm = bNorm;
            ctl->refineLast = s;
        } else {
            ctl->pcgIn.doneAt = -1;
            ctl->pcgIn.iter = 0;
            R.epsIn = epsIn;
        }
    }
    return RefTest{conv, epsIn};
}

// Solver flags of a member, written to pinned host memory after its solve
// (F_STRIDE ints per member).  F_SEQ is written last, after a system-scope
// fence, so a nonzero value tells the host the member's flags are complete.
enum { F_DONE, F_HALTED, F_ITER, F_REFDONE, F_INDONE, F_INITER, F_TT, F_SEQ, F_INNER0, F_INCONV = F_INNER0 + 8,
       F_STRIDE = 20 };

// the member's flags from its control block (one thread)
__device__ inline void write_flags(const Ctl *c, int32_t *o)
{
    o[F_DONE] = c->pcg.doneAt;
    o[F_HALTED] = c->halted;
    o[F_ITER] = c->pcg.iter;
    o[F_REFDONE] = c->ref.doneAt;
    o[F_INDONE] = c->pcgIn.doneAt;
    o[F_INITER] = c->pcgIn.iter;
    o[F_TT] = c->tt;
    for (int k = 0; k < 8; k++) o[F_INNER0 + k] = c->ref.innerIters[k];
    o[F_INCONV] = c->ref.innerConv;
    __threadfence_system();
    o[F_SEQ] = 1;
}

// The fused mixed-precision solve (k_solve_mixed): one persistent launch
// runs a group's whole free-surface solve -- refinement sweeps, inner
// iterations, the convergence tests, the surface update and the flags --
// with grid-wide barriers where separate launches had kernel boundaries.
// Its blocks must be resident together; the host sizes the grid from the
// kernel's occupancy.  Each block serves one member (plan entry: the
// member, the block's index among the member's blocks and their count) and
// walks its share of the member's tiles in every round.
struct SolveBlock {
    int member, index, count, pad;
};
// Barrier state in device memory, zero between launches: arrivals at the
// current barrier, barriers passed, blocks that had work in the round (by
// round parity), and a timeout flag (a barrier that waited longer than the
// launch's limit -- blocks not resident together -- ends the launch with
// the solve unfinished, which the host reports as an error).
struct SolveSync {
    unsigned count, gen, votes[2], err, pad[3];
};
template <int W>
__global__ void k_solve_mixed(const RefineArgs *__restrict__ R, const PcgArgs<float> *__restrict__ P,
                              const StepArgsT<float> *__restrict__ S, const SolveBlock *__restrict__ plan,
                              SolveSync *sync, int32_t *flagsOut, unsigned long long spinTicks);

// The free-surface buffers.  sf_start holds eta^n through the whole step;
// sf_new holds eta^(n-1) (the solver's extrapolated start reads it) until
// the surface update writes eta^(n+1) into it.  So the reference's surf is
// sf_start before the surface update and sf_new after it, and its surfOld
// is sf_start after it; the next step's begin flips parity, and sf_new
// becomes that step's sf_start.
template <class R>
__device__ inline gptr<double> sf_start(const StepArgsT<R> &a) { return a.ctl->parity ? a.surfB : a.surfA; }
template <class R>
__device__ inline gptr<double> sf_new(const StepArgsT<R> &a) { return a.ctl->parity ? a.surfA : a.surfB; }

template <class R>
__device__ inline gptr<const double> depth_cur(const StepArgsT<R> &a) { return a.ctl->parity ? a.depthB : a.depthA; }
template <class R>
__device__ inline gptr<double> depth_nxt(const StepArgsT<R> &a) { return a.ctl->parity ? a.depthA : a.depthB; }
// The row band of a 2-D stencil block.  Blocks are dispatched in index
// order and the last ones set a kernel's end, so the bands start at the
// tidal boundary, where the wet cells (and the subgrid lookups) and the
// boundary work sit: from the last row down unless the tide enters only
// through the YM boundary.
template <class R>
__device__ inline int row_block(const StepArgsT<R> &a)
{
    return a.rowsDown ? a.gy2 - 1 - (int)blockIdx.y : (int)blockIdx.y;
}

// the kernels after the solve run only for a member that is not halted,
// whose solve has finished and whose step they have not completed yet (a
// group launches them again when one member's solve needed more iterations)
template <class R>
__device__ inline bool tail_off(const StepArgsT<R> &a) { return a.ctl->halted || !a.ctl->runTail; }

// After a solver launch (one thread): the open flag, and, when the kernels
// after the solve are about to be launched (issue), the go-ahead for them
// -- for a member whose solve has ended and whose tail has not run in this
// step -- with the step's statistics.  (A handle starts with the go-ahead
// given, so a parity check can enter a step at its update phase.)
__device__ inline void close_solve(Ctl *c, bool mixed, bool issue)
{
    c->open = !c->halted && (mixed ? c->ref.doneAt : c->pcg.doneAt) < 0;
    if (!issue) return;
    const int run = !c->halted && !c->open && !c->tailIssued;
    c->runTail = run;
    if (!run) return;
    c->tailIssued = 1;
    c->pcgIterLast = c->pcg.iter;
    c->pcgIterTotal += c->pcg.iter;
    c->relResLast = c->pcg.bNorm > 0 ? c->pcg.rNorm / c->pcg.bNorm : 1.0;
    c->steps += 1;
    c->ttNext = c->tt + 1;
}
template <class R>
__device__ inline sptr<R> s_cur(const StepArgsT<R> &a) { return a.ctl->sParity ? a.S2 : a.S; }
template <class R>
__device__ inline sptr<R> s_nxt(const StepArgsT<R> &a) { return a.ctl->sParity ? a.S : a.S2; }

// Surface update of one index k of the Ng array: getFreeSurface +
// zeroNegSurf + enforceFreeSurfaceBC + enforceTidalBC (nsfunctions.c:
// 1020-1031, 436-449, 317-334, 412-433).  xs: the solution (x, or for the
// mixed path the refinement buffer of the sweep that converged).  A ghost
// takes the zero-clipped solution of the interior cell it mirrors (the tidal
// override comes after the ghost fill); Vloss3 (the clip) into loss.
//
// The new surface goes to the step's second buffer (sf_new below); the old
// one stays in the first (the reference's copy into surfOld).
template <class R>
__device__ inline double surface_value(const StepArgsT<R> &a, int k, int src, double &loss, gptr<const double> xs)
{
    const P &p = a.p;
    const int tt = a.ctl->tt;
    const int cell = src < 0 ? k : src;
    double h = xs[cell];
    const double zb = a.bottomZ[cell];
    if (h < zb) {
        const double diff = zb - h;
        h = zb;
        if (src < 0 && tt > 1) loss -= diff * p.dx * p.dy;
    }
    if (src < 0) {
        const uint8_t fl = a.flags[k];
        if ((fl & 1) && p.bcType != 2 && a.tideP[tt] + p.offset > zb) h = a.tideP[tt] + p.offset;
        if ((fl & 2) && p.bcType != 1 && a.tideM[tt] + p.offset > zb) h = a.tideM[tt] + p.offset;
    }
    return h;
}

template <class R>
__device__ inline void surface_cell(const StepArgsT<R> &a, int k, double &loss, gptr<const double> xs)
{
    const double h = surface_value(a, k, ghost_source(a.g, k, false), loss, xs);
    (a.ctl->parity ? a.surfA : a.surfB)[k] = h;
}

enum SubTable { T_V, T_Z, T_VXP, T_VXM, T_VYP, T_VYM, T_NP, T_NM, T_OP, T_OM };

template <class R> __global__ void k_predictor(const StepArgsT<R> *__restrict__ A);
template <class R> __global__ void k_assemble(const StepArgsT<R> *__restrict__ A);
template <class R> __global__ void k_surface(const StepArgsT<R> *__restrict__ A);
template <class R, bool FUSED> __global__ void k_limiter(const StepArgsT<R> *__restrict__ A);
template <class R> __global__ void k_faces(const StepArgsT<R> *__restrict__ A);
template <class R> __global__ void k_finish(const StepArgsT<R> *__restrict__ A);
template <class R, bool BUDGET> __global__ void k_scalar(const StepArgsT<R> *__restrict__ A);
template <class R> __global__ void k_step_end(const StepArgsT<R> *__restrict__ A);
__global__ void k_swap_surface(double *a, double *b, int n);
// subgrid variants
template <class R> __global__ void k_predictor_sub(const StepArgsT<R> *__restrict__ A);
template <class R> __global__ void k_assemble_sub(const StepArgsT<R> *__restrict__ A);
template <class R> __global__ void k_subarea_init(const StepArgsT<R> *__restrict__ A);
template <class R, bool FUSED, bool OTF> __global__ void k_limiter_sub(const StepArgsT<R> *__restrict__ A);
template <class R> __global__ void k_combine(const StepArgsT<R> *__restrict__ A);
template <class R> __global__ void k_combine_faces(const StepArgsT<R> *__restrict__ A);
template <class R> __global__ void k_finish_sub(const StepArgsT<R> *__restrict__ A);
template <class R, bool BUDGET> __global__ void k_scalar_sub(const StepArgsT<R> *__restrict__ A);

}  // namespace frehd
