// handle.h -- host-side state of the device library (internal; the public
// interface is include/frehd_gpu.h).
//
// frehd_dev   one scenario resident on the device: its arena, the host copy
//             of its kernel argument blocks and their device copies.
// Exec        a group of scenarios advanced by the same kernel launches: each
//             kernel reads its scenario's argument block from a device array
//             indexed by blockIdx.z and is sized for the group's largest
//             member (smaller members' surplus blocks exit at once).  A
//             handle alone is a group of one; frehd_batch_t is a group of
//             many.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "frehd_gpu.h"
#include "step_kernels.h"
#include "step_state.h"

namespace frehd {

int fail(int code, const char *fmt, ...);

#define HIPCHK(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) return ::frehd::fail(FREHD_EHIP, "%s failed: %s", #call,    \
                                                   hipGetErrorString(e_));                \
    } while (0)

enum KernelId { K_BEGIN, K_PRED, K_ASM, K_PQ, K_UPD, K_SURF, K_LIM, K_FACES,
                K_FINISH, K_SCALAR, K_END, K_COMBINE, K_SUBINIT,
                K_CG32, K_CG32_SET, K_CG32_STOP, K_REFINE, K_GATHER, K_PQ_STOP, K_SSOR,
                K_PCG1, K_PCG1_SET, K_PCG1_STOP, K_PCG1_START, K_SOLVE, K_COUNT,
                K_IDLE = -1 };
extern const char *kKernelNames[K_COUNT];

struct KernelTimer {
    const char *name;
    double ms;
    int64_t launches;
};

// a timed launch; solver launches carry their place in the solve (kind
// 1: FP64 iteration it, 2: inner iteration it after refinement sweep s,
// 3: refinement sweep s), classified once the flags show where each solve
// stopped: working launches keep their kernel, the one that only ran the
// stop test moves to its _stop entry, launches past convergence are dropped
struct TimedLaunch {
    int id;
    int kind, s, it;
    hipEvent_t e0, e1;
};

// per-kernel event timing on the group's stream
struct Timing {
    int on = 0;
    std::vector<KernelTimer> timers;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> pool;
    size_t next = 0;
    void reset();
    ~Timing();
};

// launch extents of a group: the maximum over its members
struct Extents {
    int gx2, gy2, nb1, nbN, nbGhost, nbRes, nbAsm, nbDir, nbRef;
    int blockX;     // the 2-D stencil kernels' block: blockX x (256 / blockX) (shared by the group)
    int lanes;      // columns per lane of k_cg1 / k_pcg1 (shared by the group)
    int ssorN;      // largest N of the members preconditioned with SSOR (0: none)
    int warm;       // an FP64 member starts its solve from a nonzero x
    int ssorLds;    // every SSOR member fits k_ssor's LDS vector
    int budget;     // a member accounts its conservation budgets
    int otf;        // a member evaluates its subgrid lookups on the fly
};

struct frehd_dev_impl;

// solver counts of one solve: FP64 iterations (n = 1, v[0]), or inner
// iterations per refinement sweep (n sweeps, at most 8); fixed storage, so
// the step's bookkeeping allocates nothing
struct SolveCounts {
    int n = 0;
    int v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

struct Exec {
    int device = 0;
    hipStream_t stream = nullptr;
    int count = 0, capacity = 0;
    StepArgs *dArgs = nullptr;          // device arrays of member argument blocks
    PcgArgs<double> *dPcg = nullptr;
    PcgArgs<float> *dPcg32 = nullptr;
    RefineArgs *dRef = nullptr;
    // F_STRIDE ints per member in pinned, coherent host memory that the
    // flag gather writes directly (hFlagsDev: its device address)
    int32_t *hFlags = nullptr, *hFlagsDev = nullptr;
    Extents ext{};
    int sub = 0, mixed = 0, nScalarMax = 0, maxIter = 0, maxRefine = 0;
    int f32 = 0;                        // the members store the explicit state in FP32 (mixed form)
    int lastIters = 16;
    SolveCounts mixPlan;                // inner iterations per sweep of the next mixed solve's pattern
    // the last steps' solver counts (FP64 iterations, or inner iterations
    // per sweep): the next pattern covers their maximum, since a launch that
    // finds its solve converged costs a few microseconds and a pattern that
    // falls short costs extra host synchronisations
    SolveCounts recent[8];
    int nRecent = 0;
    Timing *tim = nullptr;
    std::vector<frehd_dev_impl *> members;
    // captured steps (hipGraph): a whole step per solver launch pattern
    // (predictor, assembly, the pattern, the flag gather, the kernels after
    // the solve), and those last kernels alone, for a step whose solve the
    // pattern left open; rebuilt whenever the members' argument blocks change
    int useGraphs = 1;                  // launch mode: kLaunchEager, kLaunchGraphs, kLaunchStreamed, kLaunchAuto
    int patternCap = 0;                 // > 0: at most that many solver launches enqueued ahead (test hook)
    std::vector<std::pair<std::vector<int>, hipGraphExec_t>> heads;
    hipGraphExec_t tail = nullptr;
    // place in the solve of the launches being enqueued (see TimedLaunch)
    int tagKind = 0, tagS = 0, tagIt = 0;
    int pendTest = -1;                  // a refinement sweep whose outer test the next launch runs
    // capture only what repeats: steps since the last refresh, and how often
    // each launch pattern without a graph has come up
    int stableSteps = 0;
    std::vector<std::pair<std::vector<int>, int>> seen;
    // fused mixed-precision solve (k_solve_mixed): fusedMax < 0 off (the
    // default), 0 as many blocks as the device holds at once, > 0 at most
    // that many;
    // planBlocks = the grid of the current plan (0: the launched solve);
    // the plan (planCap entries) and the barrier state live in device memory
    int fusedMax = -1;
    int planBlocks = 0, planCap = 0;
    SolveBlock *dPlan = nullptr;
    SolveSync *dSync = nullptr;
    std::vector<SolveBlock> hPlan;
    unsigned long long spinTicks = 0;
};

// how a group's whole steps are launched
enum { kLaunchEager = 0, kLaunchGraphs = 1, kLaunchStreamed = 2, kLaunchAuto = 3 };

// plan entries a group's device plan array holds (blocks of one fused solve)
constexpr int kPlanCap = 4096;
// co-resident blocks of the fused solve with `lanes` columns per lane on a
// device (occupancy x compute units)
int fused_capacity(int device, int lanes);

// drop the captured steps of a group (after its members or their arguments
// change, and before it is destroyed)
void exec_release_graphs(Exec &e);

// element counts of a handle's per-block partial-sum buffers (solver:
// partA, partB, and for mixed precision partF twice and partC; volume
// losses vpart, CFL maxima cflpart) and the block counts they are sized by
struct PartialSizes {
    size_t partA, partB, vpart, cflpart, partF, partC, bpart;
    size_t maxBlocks, nb1, nb2;
};
PartialSizes partial_sizes(const frehd_config &k);

struct frehd_dev_impl {
    frehd_config cfg;
    Grid g;
    int device;
    hipStream_t stream;
    bool ownStream;
    char *arena;
    size_t arenaBytes;
    StepArgs args;
    PcgArgs<double> pcg;                // FP64 solve
    PcgArgs<float> pcg32;               // mixed precision: inner FP32 solve
    RefineArgs ref;                     // mixed precision: FP64 refinement
    int tileH, tileRemap, tileLanes;    // solver tiling as requested (tileH 0: automatic)
    int blockX;                         // 2-D stencil kernels: blocks of blockX x (256 / blockX) cells
    int midStep;                        // phases run of the step in progress (step_state.h; 0: at a step boundary)
    int tt;                             // host view of the step counter
    Exec self;                          // the handle as a group of one
    // frehd_dev_step_noalloc: device argument blocks (carved at create)
    // whose partial-sum pointers lead into the static scratch array, and
    // their host images, which stay alive while the copies are in flight
    StepArgs *naArgs;
    PcgArgs<double> *naPcg;
    PcgArgs<float> *naPcg32;
    RefineArgs *naRef;
    StepArgs naArgsH;
    PcgArgs<double> naPcgH;
    PcgArgs<float> naPcg32H;
    RefineArgs naRefH;
    Timing timing;
    void *batch;                        // the batch holding the handle, or NULL
};

// member extents
Extents member_extents(const frehd_dev_impl *d);
// (re)build a group's device argument arrays and extents from its members
int exec_refresh(Exec &e);
// run one phase (0..4) / one whole step for every member of the group
int exec_phase(Exec &e, int phase);
int exec_step(Exec &e);
// Init's subgrid part for one handle (level index, computeSubArea,
// combineSubArea)
int exec_init_subgrid(Exec &e);
int exec_collect_timing(Exec &e);
// average ms of one solver iteration under the group's current tiling
int exec_time_solver(Exec &e, int iters, double *ms);
// upload one handle's argument blocks to its own group arrays
int upload_member_args(frehd_dev_impl *d);
// the last step's CFL maxima (from the velocity update's block partials)
int cfl_maxima(frehd_dev_impl *d, double *mx, double *my);
// lets the SSOR kernel's LDS variant take `bytes` of dynamic LDS
void ssor_allow_lds(int bytes);

}  // namespace frehd

struct frehd_dev : frehd::frehd_dev_impl {};

struct frehd_batch {
    frehd::Exec e;
    frehd::Timing timing;
};
