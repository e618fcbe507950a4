// exec.hip -- kernel sequencing for a group of scenarios (Exec): the phases
// of a time step, the solver's convergence loop, the mixed-precision
// refinement loop, per-kernel timing.  Every launch covers all members of
// the group (grid z = member); see handle.h.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <hip/hip_ext.h>

#include "handle.h"

namespace frehd {

const char *kKernelNames[K_COUNT] = {"k_step_begin", "k_predictor", "k_assemble", "k_pcg_dir",
    "k_pcg_res", "k_surface", "k_limiter", "k_faces",
    "k_finish", "k_scalar", "k_step_end", "k_combine",
    "k_subarea_init", "k_cg1_f32", "k_cg1_f32_setup", "k_cg1_f32_stop", "k_refine", "k_gather_flags",
    "k_pcg_dir_stop", "k_ssor", "k_pcg1", "k_pcg1_setup", "k_pcg1_stop", "k_pcg1_start", "k_solve_mixed"};

void Timing::reset()
{
    timers.clear();
    pending.clear();
    next = 0;
}

Timing::~Timing()
{
    for (auto e : pool) hipEventDestroy(e);
}

// the solver flags of every member (one block each), gathered into one
// array so that one copy brings them to the host; a member whose solve has
// not finished is marked open, so the kernels after the solve stand down
// until it has.  testS >= 0: first the outer test of the mixed solve's
// refinement sweep testS (R: the members' refinement arguments).  out =
// nullptr: the solve is closed on the device only, no host write.
__global__ void __launch_bounds__(256) k_gather_flags(const StepArgs *__restrict__ A, int count, int mixed,
                                                      int32_t *out, const RefineArgs *__restrict__ R, int testS,
                                                      int issue)
{
    const int m = blockIdx.x;
    if (m >= count) return;
    __shared__ double lds[16 * 2];
    Ctl *c = A[m].ctl;
    // (testS is uniform, so every thread takes part in the test's reduction;
    // thread 0 records it unless the member has halted or converged before)
    if (testS >= 0) refine_test(R[m], testS, lds, !c->halted && c->ref.doneAt < 0);
    if (threadIdx.x != 0) return;
    close_solve(c, mixed != 0, issue != 0);
    if (out) write_flags(c, out + (size_t)m * F_STRIDE);
}

namespace {

hipEvent_t next_event(Timing &t)
{
    if (t.next == t.pool.size()) {
        hipEvent_t e;
        hipEventCreate(&e);
        t.pool.push_back(e);
    }
    return t.pool[t.next++];
}

// With timing on, the kernel is launched with a start and a stop event that
// the runtime stamps from the dispatch itself (hipExtLaunchKernel), so the
// elapsed time is the kernel's execution, as a kernel trace reports it --
// not the gap between two event records around the launch.
template <class F>
int launch(Exec &e, KernelId id, F &&fn)
{
    Timing *t = e.tim;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (t && t->on) {
        e0 = next_event(*t);
        e1 = next_event(*t);
    }
    fn(e0, e1);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return fail(FREHD_EHIP, "launch of %s failed: %s", kKernelNames[id], hipGetErrorString(err));
    if (e0) t->pending.push_back({id, e.tagKind, e.tagS, e.tagIt, e0, e1});
    return 0;
}

#define LAUNCH_SHM(id, kern, grid, block, shm, ...)                                                    \
    do {                                                                                              \
        int rc_ = launch(e, id, [&](hipEvent_t a_, hipEvent_t b_) {                                   \
            if (a_) hipExtLaunchKernelGGL(kern, grid, block, shm, e.stream, a_, b_, 0, __VA_ARGS__);  \
            else hipLaunchKernelGGL(kern, grid, block, shm, e.stream, __VA_ARGS__);                   \
        });                                                                                           \
        if (rc_) return rc_;                                                                          \
    } while (0)
#define LAUNCH(id, kern, grid, block, ...) LAUNCH_SHM(id, kern, grid, block, 0, __VA_ARGS__)
// a step kernel in the group's state precision: FP32 storage for the mixed
// form, FP64 otherwise (the argument blocks are StepArgs in both; the FP32
// instantiation reads the same pointers as StepArgsT<float>)
#define LAUNCH_ST(id, kern, grid, block)                                                               \
    do {                                                                                              \
        if (e.f32) LAUNCH(id, kern<float>, grid, block, reinterpret_cast<const StepArgsT<float> *>(e.dArgs)); \
        else LAUNCH(id, kern<double>, grid, block, e.dArgs);                                          \
    } while (0)

inline dim3 grid1(int nb, int n) { return dim3((unsigned)std::max(nb, 1), 1, (unsigned)n); }

// keep the solver counts of the last 8 solves
void remember(Exec &e, const SolveCounts &counts)
{
    if (e.nRecent == 8) {
        for (int q = 1; q < 8; q++) e.recent[q - 1] = e.recent[q];
        e.nRecent = 7;
    }
    e.recent[e.nRecent++] = counts;
}

// issue: the kernels after the solve follow this gather (members whose
// solve has ended get their go-ahead).  host: the flags go to the host
// buffer too.  A gather the host does not wait for must not write them: it
// could land after the host has cleared the flags for the next step, whose
// wait would then take them for that step's (and miss an open solve).
int enqueue_gather(Exec &e, int issue, bool host)
{
    const int testS = e.mixed ? e.pendTest : -1;
    e.pendTest = -1;
    LAUNCH(K_GATHER, k_gather_flags, dim3((unsigned)e.count), dim3(256), e.dArgs, e.count, e.mixed,
           host ? e.hFlagsDev : (int32_t *)nullptr, e.dRef, testS, issue);
    return 0;
}

// wait until the gather of every member has written its flags (the host
// cleared F_SEQ before the launch), checking the stream for errors now and
// then; the stream keeps running meanwhile
int wait_flags(Exec &e)
{
    volatile int32_t *f = e.hFlags;
    for (long spin = 0;; spin++) {
        bool all = true;
        for (int m = 0; m < e.count && all; m++)
            if (f[(size_t)m * F_STRIDE + F_SEQ] == 0) all = false;
        if (all) break;
        if ((spin & 1023) == 1023) {
            const hipError_t q = hipStreamQuery(e.stream);
            if (q != hipSuccess && q != hipErrorNotReady)
                return fail(FREHD_EHIP, "step failed: %s", hipGetErrorString(q));
            if (q == hipSuccess) {
                // the stream drained: the flags are there unless the gather
                // did not run, which is an error
                for (int m = 0; m < e.count; m++)
                    if (f[(size_t)m * F_STRIDE + F_SEQ] == 0) return fail(FREHD_EHIP, "solver flags missing");
                break;
            }
        }
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return 0;
}

void clear_flags(Exec &e)
{
    volatile int32_t *f = e.hFlags;
    for (int m = 0; m < e.count; m++) f[(size_t)m * F_STRIDE + F_SEQ] = 0;
    std::atomic_thread_fence(std::memory_order_seq_cst);
}

int gather(Exec &e)
{
    int rc = enqueue_gather(e, 0, true);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e.stream));
    return 0;
}

inline const int32_t *flags(const Exec &e, int m) { return e.hFlags + (size_t)m * F_STRIDE; }

struct Tag {
    Exec &e;
    Tag(Exec &ex, int kind, int s, int it) : e(ex) { e.tagKind = kind; e.tagS = s; e.tagIt = it; }
    ~Tag() { e.tagKind = 0; }
};

// launch it of the FP64 solve: with Jacobi one k_pcg1 (it = 0 sets up,
// after k_pcg1_start for members starting from a nonzero x); with SSOR
// (a group's members share the preconditioner) k_ssor, k_pcg_dir, k_pcg_res
int enqueue_pcg_iteration(Exec &e, int it)
{
    Tag tag(e, 1, 0, it);
    if (e.ext.ssorN > 0) {
        // z = M^-1 r of the members preconditioned with SSOR (one workgroup each)
        const dim3 g(1, 1, (unsigned)e.count), b(1024);
        if (e.ext.ssorLds) LAUNCH_SHM(K_SSOR, k_ssor<true>, g, b, (size_t)e.ext.ssorN * sizeof(double), e.dPcg, it);
        else LAUNCH(K_SSOR, k_ssor<false>, g, b, e.dPcg, it);
        LAUNCH(K_PQ, k_pcg_dir<double>, grid1(e.ext.nbDir, e.count), dim3(256), e.dPcg, it);
        LAUNCH(K_UPD, k_pcg_res<double>, grid1(e.ext.nbRes, e.count), dim3(256), e.dPcg, it);
        return 0;
    }
    const dim3 g = grid1(e.ext.nbDir, e.count), b(256);
    if (it == 0 && e.ext.warm) LAUNCH(K_PCG1_START, k_pcg1_start, grid1(e.ext.nbRes, e.count), b, e.dPcg);
    switch ((it == 0 ? 8 : 0) + e.ext.lanes) {
    case 9: LAUNCH(K_PCG1_SET, (k_pcg1<true, 1>), g, b, e.dPcg, it); break;
    case 10: LAUNCH(K_PCG1_SET, (k_pcg1<true, 2>), g, b, e.dPcg, it); break;
    case 1: LAUNCH(K_PCG1, (k_pcg1<false, 1>), g, b, e.dPcg, it); break;
    case 2: LAUNCH(K_PCG1, (k_pcg1<false, 2>), g, b, e.dPcg, it); break;
    default: return fail(FREHD_EINVAL, "columns per lane %d", e.ext.lanes);
    }
    return 0;
}

// launches of an FP64 solve of n iterations, the one that stops included:
// k_pcg1 launches 0..n+1 (launch L tests the residual of iteration L-1),
// SSOR iterations 0..n (k_pcg_dir(n) tests the residual of iteration n)
inline int fp64_launches(const Exec &e, int n) { return e.ext.ssorN > 0 ? n + 1 : n + 2; }

// After a solve, the timed solver launches by what they did (see
// TimedLaunch): FP64 launch it works below the slowest member's stop
// launch D, which only tests (k_pcg_dir(D) also applies the last x update
// of the SSOR path); k_pcg1_start always works; inner
// iteration it after sweep s works up to the most iterations any member
// used there, the next launch only stops; a sweep works up to the slowest
// member's converged sweep.
void classify_timed(Exec &e)
{
    Timing *t = e.tim;
    if (!t || !t->on) return;
    int D = 0, R = 0;
    for (int m = 0; m < e.count; m++) {
        if (flags(e, m)[F_HALTED]) continue;
        D = std::max(D, (int)flags(e, m)[F_DONE]);
        R = std::max(R, (int)flags(e, m)[F_REFDONE]);
    }
    auto inner = [&](int s) {
        int n = -1;
        for (int m = 0; m < e.count; m++)
            if (!flags(e, m)[F_HALTED] && flags(e, m)[F_REFDONE] > s && s < 8)
                n = std::max(n, (int)flags(e, m)[F_INNER0 + s]);
        return n;   // -1: no member solved after sweep s
    };
    for (auto &r : t->pending) {
        if (r.kind == 1) {
            if (r.id == K_SSOR) r.id = r.it <= D ? K_SSOR : K_IDLE;
            else if (r.id == K_PQ) r.id = r.it < D ? K_PQ : (r.it == D ? K_PQ_STOP : K_IDLE);
            else if (r.id == K_UPD) r.id = r.it < D ? K_UPD : K_IDLE;
            else if (r.id == K_PCG1 || r.id == K_PCG1_SET) r.id = r.it < D ? r.id : (r.it == D ? K_PCG1_STOP : K_IDLE);
        } else if (r.kind == 2) {
            const int n = inner(r.s);
            if (n < 0) r.id = K_IDLE;
            else if (r.it == 0) r.id = K_CG32_SET;
            else r.id = r.it <= n ? K_CG32 : (r.it == n + 1 ? K_CG32_STOP : K_IDLE);
        } else if (r.kind == 3) {
            if (r.s > R) r.id = K_IDLE;
        }
        r.kind = 0;
    }
}

// FP64 PCG.  The step enqueues the iteration count that covers the last
// solves and gathers the flags; solve_fp64_finish then adds small chunks
// until every member reports convergence (or a halt).
int solve_fp64_pattern(Exec &e, int *it)
{
    int n = fp64_launches(e, std::max(1, e.lastIters));
    if (e.patternCap > 0) n = std::min(n, e.patternCap);
    for (int q = 0; q < n; q++) {
        int rc = enqueue_pcg_iteration(e, q);
        if (rc) return rc;
    }
    *it = n;
    return 0;
}

int solve_fp64_finish(Exec &e, int it)
{
    for (;;) {
        bool all = true;
        for (int m = 0; m < e.count; m++)
            if (flags(e, m)[F_DONE] < 0 && !flags(e, m)[F_HALTED]) all = false;
        if (all) break;
        if (it > e.maxIter + 8) return fail(FREHD_EHIP, "PCG did not report completion");
        for (int q = 0; q < 4; q++, it++) {
            int rc = enqueue_pcg_iteration(e, it);
            if (rc) return rc;
        }
        int rc = gather(e);
        if (rc) return rc;
    }
    classify_timed(e);
    int last = 1;
    for (int m = 0; m < e.count; m++)
        if (!flags(e, m)[F_HALTED]) last = std::max(last, (int)flags(e, m)[F_ITER]);
    SolveCounts c;
    c.n = 1;
    c.v[0] = last;
    remember(e, c);
    e.lastIters = 1;
    for (int q = 0; q < e.nRecent; q++) e.lastIters = std::max(e.lastIters, e.recent[q].v[0]);
    return 0;
}

// launch it of the inner FP32 solve after refinement sweep s (iteration
// it; 0 sets up, and runs the outer test of the sweep before it when one is
// pending)
int enqueue_inner_iteration(Exec &e, int s, int it)
{
    Tag tag(e, 2, s, it);
    const dim3 g = grid1(e.ext.nbDir, e.count), b(256);
    const int testS = it == 0 ? e.pendTest : -1;
    if (it == 0) e.pendTest = -1;
    const RefineArgs *R = e.dRef;
    switch ((it == 0 ? 8 : 0) + e.ext.lanes) {
    case 9: LAUNCH(K_CG32_SET, (k_cg1<true, 1>), g, b, e.dPcg32, it, R, testS); break;
    case 10: LAUNCH(K_CG32_SET, (k_cg1<true, 2>), g, b, e.dPcg32, it, R, testS); break;
    case 12: LAUNCH(K_CG32_SET, (k_cg1<true, 4>), g, b, e.dPcg32, it, R, testS); break;
    case 1: LAUNCH(K_CG32, (k_cg1<false, 1>), g, b, e.dPcg32, it, R, -1); break;
    case 2: LAUNCH(K_CG32, (k_cg1<false, 2>), g, b, e.dPcg32, it, R, -1); break;
    case 4: LAUNCH(K_CG32, (k_cg1<false, 4>), g, b, e.dPcg32, it, R, -1); break;
    default: return fail(FREHD_EINVAL, "columns per lane %d", e.ext.lanes);
    }
    return 0;
}

// refinement sweep s; its outer test runs in the next launch (the inner
// solve's set-up or the flag gather), which takes it from pendTest
int enqueue_refine(Exec &e, int s)
{
    Tag tag(e, 3, s, 0);
    LAUNCH(K_REFINE, k_refine, grid1(e.ext.nbRef, e.count), dim3(256), e.dRef, s, e.dPcg32);
    e.pendTest = s;
    return 0;
}

// one inner solve of n iterations: the set-up launch and n iterations
// (launches past a member's convergence exit at once); the stop test on the
// residual of iteration n is made by the refinement sweep after it
int enqueue_inner_solve(Exec &e, int s, int n)
{
    for (int it = 0; it <= n; it++) {
        int rc = enqueue_inner_iteration(e, s, it);
        if (rc) return rc;
    }
    return 0;
}

// Mixed-precision solve: sweeps of FP64 refinement around FP32 inner
// solves.  The launch pattern of the last solve (sweeps and inner
// iterations per sweep, the maximum over the members) is enqueued without
// waiting; launches past a member's convergence exit at once, and an inner
// solve cut short by the pattern still yields a valid correction (the next
// sweep recomputes the true residual).  If the pattern ends before every
// member's outer test passes, further sweeps follow, each inner solve run to
// its own stop test.
int solve_mixed_pattern(Exec &e, int *sweeps)
{
    int rc, s = 0;
    e.pendTest = -1;
    if ((rc = enqueue_refine(e, 0))) return rc;
    const int nk = e.patternCap > 0 ? std::min(e.mixPlan.n, 1) : e.mixPlan.n;
    for (int k = 0; k < nk; k++) {
        const int n = e.patternCap > 0 ? std::min(e.mixPlan.v[k], e.patternCap) : e.mixPlan.v[k];
        if ((rc = enqueue_inner_solve(e, s, n))) return rc;
        if ((rc = enqueue_refine(e, ++s))) return rc;
    }
    *sweeps = s;
    return 0;
}

int solve_mixed_finish(Exec &e, int s)
{
    int rc;
    auto outer_open = [&](int m) { return flags(e, m)[F_REFDONE] < 0 && !flags(e, m)[F_HALTED]; };
    auto any_open = [&] {
        for (int m = 0; m < e.count; m++)
            if (outer_open(m)) return true;
        return false;
    };
    while (any_open()) {
        if (s >= e.maxRefine) return fail(FREHD_EHIP, "refinement did not report completion");
        int it = 0;
        for (;;) {
            for (int q = 0; q < 8; q++, it++)
                if ((rc = enqueue_inner_iteration(e, s, it))) return rc;
            if ((rc = gather(e))) return rc;
            bool all = true;
            for (int m = 0; m < e.count; m++)
                if (outer_open(m) && flags(e, m)[F_INDONE] < 0) all = false;
            if (all) break;
            if (it > e.maxIter + 8) return fail(FREHD_EHIP, "inner PCG did not report completion");
        }
        if ((rc = enqueue_refine(e, ++s))) return rc;
        if ((rc = gather(e))) return rc;
    }
    classify_timed(e);
    // this solve's counts: per sweep, the most inner iterations any member
    // used, two more where the pattern cut a member's inner solve short (it
    // ended before its own stop test passed) and the member then needed
    // another sweep; the next pattern covers the last steps' counts sweep by
    // sweep
    int sweeps = 0;
    for (int m = 0; m < e.count; m++)
        if (!flags(e, m)[F_HALTED]) sweeps = std::max(sweeps, (int)flags(e, m)[F_REFDONE]);
    SolveCounts plan;
    for (int k = 0; k < sweeps && k < 8; k++) {
        int n = 1;
        bool cut = false;
        for (int m = 0; m < e.count; m++) {
            const int32_t *f = flags(e, m);
            if (f[F_HALTED] || f[F_REFDONE] <= k) continue;
            n = std::max(n, (int)f[F_INNER0 + k]);
            if (!((f[F_INCONV] >> k) & 1) && f[F_REFDONE] > k + 1) cut = true;
        }
        if (k < e.mixPlan.n && cut && n >= e.mixPlan.v[k]) n = e.mixPlan.v[k] + 2;
        plan.v[plan.n++] = n;
    }
    remember(e, plan);
    e.mixPlan = SolveCounts();
    for (int q = 0; q < e.nRecent; q++) {
        const SolveCounts &r = e.recent[q];
        e.mixPlan.n = std::max(e.mixPlan.n, r.n);
        for (int k = 0; k < r.n; k++) e.mixPlan.v[k] = std::max(e.mixPlan.v[k], r.v[k]);
    }
    return 0;
}

// The whole mixed-precision solve of every member, its surface update and
// its flags in one persistent launch (k_solve_mixed; the group's plan gives
// each block its member and share).
int enqueue_fused_solve(Exec &e)
{
    const auto *a32 = reinterpret_cast<const StepArgsT<float> *>(e.dArgs);
    const dim3 g((unsigned)e.planBlocks), b(256);
    switch (e.ext.lanes) {
    case 1: LAUNCH(K_SOLVE, (k_solve_mixed<1>), g, b, e.dRef, e.dPcg32, a32, e.dPlan, e.dSync, e.hFlagsDev, e.spinTicks); break;
    case 2: LAUNCH(K_SOLVE, (k_solve_mixed<2>), g, b, e.dRef, e.dPcg32, a32, e.dPlan, e.dSync, e.hFlagsDev, e.spinTicks); break;
    case 4: LAUNCH(K_SOLVE, (k_solve_mixed<4>), g, b, e.dRef, e.dPcg32, a32, e.dPlan, e.dSync, e.hFlagsDev, e.spinTicks); break;
    default: return fail(FREHD_EINVAL, "columns per lane %d", e.ext.lanes);
    }
    return 0;
}

// A fused solve that did not deliver every member's flags: its barrier
// timed out (the blocks were not resident together) or it did not finish.
// The barrier state is cleared for the next launch; the step fails.
int fused_failed(Exec &e, int rc)
{
    (void)hipStreamSynchronize(e.stream);
    SolveSync st{};
    (void)hipMemcpy(&st, e.dSync, sizeof(SolveSync), hipMemcpyDeviceToHost);
    (void)hipMemsetAsync(e.dSync, 0, sizeof(SolveSync), e.stream);
    (void)hipStreamSynchronize(e.stream);
    if (st.err)
        return fail(FREHD_EHIP, "fused solve: a grid barrier timed out (its %d blocks were not resident together; "
                                "run fewer fused solves at once, or switch it off)", e.planBlocks);
    return rc ? rc : fail(FREHD_EHIP, "fused solve: a member's solve did not finish");
}

// after a fused solve: every member's flags arrived and every solve that
// ran converged
int fused_check(Exec &e)
{
    int rc = wait_flags(e);
    if (rc) return fused_failed(e, rc);
    for (int m = 0; m < e.count; m++)
        if (!flags(e, m)[F_HALTED] && flags(e, m)[F_REFDONE] < 0) return fused_failed(e, 0);
    return 0;
}

}  // namespace

// Average duration (ms) of one solver iteration of the group under its
// current tiling: `iters` iterations on the members' current systems, timed
// with events on the group's stream (FP64: one k_pcg1 launch, or the SSOR
// triple k_ssor + k_pcg_dir + k_pcg_res;
// mixed: the inner k_cg1 launches after k_refine(0) launches, the
// refinement sweeps spread over the inner iterations in the proportion of
// the last solve).  The caller makes the iterations run (zero tolerance)
// and restores the control blocks.
int exec_time_solver(Exec &e, int iters, double *ms)
{
    Timing local;
    local.on = 1;
    struct Swap {
        Exec &e;
        Timing *saved;
        ~Swap() { e.tim = saved; }
    } swap{e, e.tim};
    e.tim = &local;
    e.pendTest = -1;
    int rc;
    if (e.mixed) {
        for (int q = 0; q < 4; q++)
            LAUNCH(K_REFINE, k_refine, grid1(e.ext.nbRef, e.count), dim3(256), e.dRef, 0, e.dPcg32);
        for (int it = 0; it <= iters; it++)
            if ((rc = enqueue_inner_iteration(e, 0, it))) return rc;
    } else {
        // each trial solves from the step's start residual r = b (earlier
        // trials have iterated r down to rounding noise, on which the
        // directions can break down and the launches exit at once); a warm
        // member's k_pcg1_start recomputes it from b
        for (auto *d : e.members)
            if (!d->pcg.warm)
                HIPCHK(hipMemcpyAsync(d->args.r, d->args.b, (size_t)d->g.N * sizeof(double), hipMemcpyDeviceToDevice,
                                      e.stream));
        for (int it = 0; it <= iters; it++)
            if ((rc = enqueue_pcg_iteration(e, it))) return rc;
    }
    if ((rc = exec_collect_timing(e))) return rc;
    auto avg = [&](int k) {
        const KernelTimer &t = local.timers[k];
        return t.launches ? t.ms / (double)t.launches : 0.0;
    };
    if (e.mixed) {
        // per solve as the last step ran it: its inner iterations and one
        // refinement sweep more than inner solves
        int inner = 0;
        for (int k = 0; k < e.mixPlan.n; k++) inner += e.mixPlan.v[k];
        const int sweeps = e.mixPlan.n + 1;
        *ms = (std::max(inner, 1) * avg(K_CG32) + sweeps * avg(K_REFINE)) / std::max(inner, 1);
    } else if (e.ext.ssorN > 0) {
        *ms = avg(K_SSOR) + avg(K_PQ) + avg(K_UPD);
    } else {
        *ms = avg(K_PCG1);
    }
    return 0;
}

int exec_collect_timing(Exec &e)
{
    Timing *t = e.tim;
    if (!t || !t->on) return 0;
    HIPCHK(hipStreamSynchronize(e.stream));
    if (t->timers.empty()) {
        t->timers.resize(K_COUNT);
        for (int q = 0; q < K_COUNT; q++) t->timers[q] = KernelTimer{kKernelNames[q], 0.0, 0};
    }
    for (auto &pe : t->pending) {
        if (pe.id == K_IDLE) continue;
        float ms = 0;
        hipEventElapsedTime(&ms, pe.e0, pe.e1);
        t->timers[pe.id].ms += ms;
        t->timers[pe.id].launches += 1;
    }
    t->pending.clear();
    t->next = 0;
    return 0;
}

int enqueue_assemble(Exec &e)
{
    const int n = e.count;
    if (e.sub) LAUNCH_ST(K_ASM, k_assemble_sub, grid1(e.ext.nbAsm, n), dim3(256));
    else LAUNCH_ST(K_ASM, k_assemble, grid1(e.ext.nbAsm, n), dim3(256));
    return 0;
}

static int exec_phase_launch(Exec &e, int phase);

// updateData (+ the CFL numbers of monitorCFL); fusedSurface: the limiter
// also makes the surface update (no separate surface sweep before it)
static int enqueue_update(Exec &e, bool fusedSurface)
{
    const int n = e.count;
    const Extents &x = e.ext;
    const dim3 b2((unsigned)x.blockX, (unsigned)(256 / x.blockX)), g2((unsigned)x.gx2, (unsigned)x.gy2, (unsigned)n),
        b1(256);
    const auto *a32 = reinterpret_cast<const StepArgsT<float> *>(e.dArgs);
    const int v = (e.f32 ? 2 : 0) + (fusedSurface ? 1 : 0);
    if (e.sub) {
        switch (v + (x.otf ? 4 : 0)) {
        case 0: LAUNCH(K_LIM, (k_limiter_sub<double, false, false>), g2, b2, e.dArgs); break;
        case 1: LAUNCH(K_LIM, (k_limiter_sub<double, true, false>), g2, b2, e.dArgs); break;
        case 2: LAUNCH(K_LIM, (k_limiter_sub<float, false, false>), g2, b2, a32); break;
        case 3: LAUNCH(K_LIM, (k_limiter_sub<float, true, false>), g2, b2, a32); break;
        case 4: LAUNCH(K_LIM, (k_limiter_sub<double, false, true>), g2, b2, e.dArgs); break;
        case 5: LAUNCH(K_LIM, (k_limiter_sub<double, true, true>), g2, b2, e.dArgs); break;
        case 6: LAUNCH(K_LIM, (k_limiter_sub<float, false, true>), g2, b2, a32); break;
        default: LAUNCH(K_LIM, (k_limiter_sub<float, true, true>), g2, b2, a32); break;
        }
        LAUNCH_ST(K_FACES, k_combine_faces, grid1(x.nb1, n), b1);
        LAUNCH_ST(K_FINISH, k_finish_sub, grid1(x.nb1, n), b1);
    } else {
        switch (v) {
        case 0: LAUNCH(K_LIM, (k_limiter<double, false>), g2, b2, e.dArgs); break;
        case 1: LAUNCH(K_LIM, (k_limiter<double, true>), g2, b2, e.dArgs); break;
        case 2: LAUNCH(K_LIM, (k_limiter<float, false>), g2, b2, a32); break;
        default: LAUNCH(K_LIM, (k_limiter<float, true>), g2, b2, a32); break;
        }
        LAUNCH_ST(K_FACES, k_faces, grid1(x.nb1, n), b1);
        LAUNCH_ST(K_FINISH, k_finish, grid1(x.nb1, n), b1);
    }
    return 0;
}

int exec_phase(Exec &e, int phase)
{
    int rc = exec_phase_launch(e, phase);
    if (rc) return rc;
    // where the members are inside the step (downloads between phases)
    for (auto *m : e.members) m->midStep = step_state_after(phase);
    return 0;
}

static int exec_phase_launch(Exec &e, int phase)
{
    const int n = e.count;
    if (n == 0) return 0;
    const Extents &x = e.ext;
    const dim3 b2((unsigned)x.blockX, (unsigned)(256 / x.blockX)), g2((unsigned)x.gx2, (unsigned)x.gy2, (unsigned)n),
        b1(256);
    switch (phase) {
    case 0:   // solveSourceTerm (up to the RHS, which the assembly computes)
        if (e.sub) LAUNCH_ST(K_PRED, k_predictor_sub, g2, b2);
        else LAUNCH_ST(K_PRED, k_predictor, g2, b2);
        return 0;
    case 1: { // solveFreeSurface
        int rc = enqueue_assemble(e), k = 0;
        if (!rc && e.planBlocks > 0) {
            // the fused solve updates the surface too
            clear_flags(e);
            if ((rc = enqueue_fused_solve(e))) return rc;
            return fused_check(e);
        }
        if (!rc) rc = e.mixed ? solve_mixed_pattern(e, &k) : solve_fp64_pattern(e, &k);
        if (!rc) rc = gather(e);
        if (!rc) rc = e.mixed ? solve_mixed_finish(e, k) : solve_fp64_finish(e, k);
        if (!rc) rc = enqueue_gather(e, 1, false);   // (the flags came with the finish)
        if (rc) return rc;
        LAUNCH_ST(K_SURF, k_surface, grid1(x.nb1, n), b1);
        return 0;
    }
    case 2:   // updateData (+ the CFL numbers of monitorCFL)
        return enqueue_update(e, false);
    case 3:   // scalarTransport, every scalar in one launch (members without scalars skip inside)
        if (e.nScalarMax > 0) {
            const auto *a32 = reinterpret_cast<const StepArgsT<float> *>(e.dArgs);
            const int v = (e.sub ? 4 : 0) + (e.f32 ? 2 : 0) + (e.ext.budget ? 1 : 0);
            switch (v) {
            case 0: LAUNCH(K_SCALAR, (k_scalar<double, false>), g2, b2, e.dArgs); break;
            case 1: LAUNCH(K_SCALAR, (k_scalar<double, true>), g2, b2, e.dArgs); break;
            case 2: LAUNCH(K_SCALAR, (k_scalar<float, false>), g2, b2, a32); break;
            case 3: LAUNCH(K_SCALAR, (k_scalar<float, true>), g2, b2, a32); break;
            case 4: LAUNCH(K_SCALAR, (k_scalar_sub<double, false>), g2, b2, e.dArgs); break;
            case 5: LAUNCH(K_SCALAR, (k_scalar_sub<double, true>), g2, b2, e.dArgs); break;
            case 6: LAUNCH(K_SCALAR, (k_scalar_sub<float, false>), g2, b2, a32); break;
            default: LAUNCH(K_SCALAR, (k_scalar_sub<float, true>), g2, b2, a32); break;
            }
        }
        return 0;
    case 4:   // monitorCFL: its stop test ran in the velocity update; the budgets' step end
        if (e.ext.budget) LAUNCH_ST(K_END, k_step_end, dim3(1, 1, n), dim3(1024));
        for (auto *m : e.members) m->tt += 1;
        return 0;
    default:
        return fail(FREHD_EINVAL, "phase %d out of range", phase);
    }
}

// A step returns while the tail of its graph still runs (the next step's
// graph queues behind it), so the stream drains before any executable
// graph is destroyed: the runtime frees a graph's argument buffers with it.
void exec_release_graphs(Exec &e)
{
    if ((!e.heads.empty() || e.tail) && e.stream) hipStreamSynchronize(e.stream);
    for (auto &h : e.heads) hipGraphExecDestroy(h.second);
    e.heads.clear();
    if (e.tail) hipGraphExecDestroy(e.tail);
    e.tail = nullptr;
    e.stableSteps = 0;
    e.seen.clear();
}

// capture what `enqueue` launches on the group's stream into an executable
// graph
template <class F>
static int capture(Exec &e, F &&enqueue, hipGraphExec_t *out)
{
    HIPCHK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue();
    hipGraph_t g = nullptr;
    hipError_t err = hipStreamEndCapture(e.stream, &g);
    if (rc || err != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return rc ? rc : fail(FREHD_EHIP, "stream capture failed: %s", hipGetErrorString(err));
    }
    err = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (err != hipSuccess) return fail(FREHD_EHIP, "graph instantiation failed: %s", hipGetErrorString(err));
    return 0;
}

static int exec_step_eager(Exec &e)
{
    for (int ph = 0; ph < 5; ph++) {
        int rc = exec_phase(e, ph);
        if (rc) return rc;
    }
    return 0;
}

// enqueue the kernels after the solve in a captured step: the update --
// with the surface update fused into its limiter, unless the fused solve
// wrote the surface --, the scalars, the step end
static int enqueue_tail(Exec &e)
{
    int r = enqueue_update(e, e.planBlocks == 0);
    if (r) return r;
    for (auto *m : e.members) m->midStep = step_state_after(2);
    for (int ph = 3; ph <= 4; ph++)
        if ((r = exec_phase(e, ph))) return r;
    return 0;
}

// One step as one graph launch per launch pattern: the predictor, the
// assembly, the solver's pattern, the flag gather -- which writes the flags
// straight to host memory -- and the kernels after the solve, which stand
// down for a member the gather marked open.  The host polls the flags while
// the rest of the graph runs and, normally, returns, so the next step's
// graph is queued behind this one and the GPU does not idle between steps.
// If a solve is still open, the host waits for the graph, finishes the solve
// with eager iterations (their gather closes it) and launches the tail
// (captured on its own) for that member; members whose tail ran stand down.
// The same kernels with the same arguments as the eager path, so the results
// are identical.
// One step as plain launches with the same structure as a captured step:
// the predictor, the assembly, the solver's pattern and the flag gather,
// then -- without waiting -- the kernels after the solve, which stand down
// for a member whose solve the pattern left open; the host then waits for
// the flags (while the tail runs) and finishes any open solve as the
// captured step does.  The host queues a step's launches well ahead of the
// device, and consecutive launches on a stream follow each other more
// closely than consecutive graph launches (about 14 us lie between the end
// of one step graph and the start of the next on MI355X), so large grids
// step faster this way.
static int exec_step_streamed(Exec &e)
{
    const int n = e.count;
    int k = 0, rc;
    clear_flags(e);
    {
        const dim3 b2((unsigned)e.ext.blockX, (unsigned)(256 / e.ext.blockX)),
            g2((unsigned)e.ext.gx2, (unsigned)e.ext.gy2, (unsigned)n);
        if (e.sub) LAUNCH_ST(K_PRED, k_predictor_sub, g2, b2);
        else LAUNCH_ST(K_PRED, k_predictor, g2, b2);
        int r = enqueue_assemble(e);
        if (e.planBlocks > 0) {
            if (!r) r = enqueue_fused_solve(e);
        } else {
            if (!r) r = e.mixed ? solve_mixed_pattern(e, &k) : solve_fp64_pattern(e, &k);
            if (!r) r = enqueue_gather(e, 1, true);
        }
        if (!r) r = enqueue_tail(e);     // (its phase-4 bookkeeping advances the members' step)
        if (r) return r;
    }
    if (e.planBlocks > 0) return fused_check(e);
    if ((rc = wait_flags(e))) return rc;
    bool open = false;
    for (int m = 0; m < n; m++)
        if (!flags(e, m)[F_HALTED] && flags(e, m)[e.mixed ? F_REFDONE : F_DONE] < 0) open = true;
    // finish the open solves (behind the stood-down tail) and record the
    // counts for the next pattern
    rc = e.mixed ? solve_mixed_finish(e, k) : solve_fp64_finish(e, k);
    if (rc) return rc;
    if (open) {
        if ((rc = enqueue_gather(e, 1, false))) return rc;   // (not waited for)
        if ((rc = enqueue_tail(e))) return rc;
        for (auto *m : e.members) m->tt -= 1;   // the step advanced once
    }
    return 0;
}

static int exec_step_graphs(Exec &e)
{
    const int n = e.count;
    std::vector<int> key;
    if (e.planBlocks > 0) {
        key = {2};
    } else if (e.mixed) {
        key.push_back(1);
        key.insert(key.end(), e.mixPlan.v, e.mixPlan.v + e.mixPlan.n);
    } else {
        key = {0, std::max(1, e.lastIters)};
    }
    hipGraphExec_t step = nullptr;
    for (auto &h : e.heads)
        if (h.first == key) step = h.second;
    int k = 0, rc;
    if (!step) {
        // a pattern is captured the second time it comes up; the first time
        // the step streams its launches
        int *count = nullptr;
        for (auto &sk : e.seen)
            if (sk.first == key) count = &sk.second;
        if (!count) {
            if (e.seen.size() >= 64) e.seen.clear();
            e.seen.push_back({key, 0});
            count = &e.seen.back().second;
        }
        if (++*count < 2) return exec_step_streamed(e);
        if (e.heads.size() >= 16) exec_release_graphs(e);
        if (getenv("FREHD_GRAPH_LOG")) {
            fprintf(stderr, "frehd: capturing step graph %zu, pattern", e.heads.size());
            for (int v : key) fprintf(stderr, " %d", v);
            fprintf(stderr, " (member 0 at step %d)\n", e.count ? e.members[0]->tt + 1 : -1);
        }
        rc = capture(e, [&] {
            const dim3 b2((unsigned)e.ext.blockX, (unsigned)(256 / e.ext.blockX)),
                g2((unsigned)e.ext.gx2, (unsigned)e.ext.gy2, (unsigned)n);
            if (e.sub) LAUNCH_ST(K_PRED, k_predictor_sub, g2, b2);
            else LAUNCH_ST(K_PRED, k_predictor, g2, b2);
            int r = enqueue_assemble(e);
            if (e.planBlocks > 0) {
                if (!r) r = enqueue_fused_solve(e);
            } else {
                if (!r) r = e.mixed ? solve_mixed_pattern(e, &k) : solve_fp64_pattern(e, &k);
                if (!r) r = enqueue_gather(e, 1, true);
            }
            if (!r) r = enqueue_tail(e);
            return r;
        }, &step);
        if (rc) return rc;
        // the capture ran the host bookkeeping of phase 4 once
        for (auto *m : e.members) m->tt -= 1;
        e.heads.push_back({key, step});
    } else {
        // (what the captured pattern enqueued: its sweeps, or its FP64 launches)
        k = e.mixed ? (e.patternCap > 0 ? std::min(e.mixPlan.n, 1) : e.mixPlan.n)
                    : fp64_launches(e, std::max(1, e.lastIters));
        if (!e.mixed && e.patternCap > 0) k = std::min(k, e.patternCap);
    }
    clear_flags(e);
    HIPCHK(hipGraphLaunch(step, e.stream));
    if (e.planBlocks > 0) {
        // the fused solve finishes every solve on the device
        if ((rc = fused_check(e))) return rc;
        for (auto *m : e.members) m->tt += 1;
        return 0;
    }
    if ((rc = wait_flags(e))) return rc;
    bool open = false;
    for (int m = 0; m < n; m++)
        if (!flags(e, m)[F_HALTED] && flags(e, m)[e.mixed ? F_REFDONE : F_DONE] < 0) open = true;
    // finish the open solves (behind the stood-down tail) and record the
    // counts for the next pattern
    rc = e.mixed ? solve_mixed_finish(e, k) : solve_fp64_finish(e, k);
    if (rc) return rc;
    if (open) {
        if ((rc = enqueue_gather(e, 1, false))) return rc;   // (not waited for)
        if (!e.tail) {
            rc = capture(e, [&] { return enqueue_tail(e); }, &e.tail);
            if (rc) return rc;
            for (auto *m : e.members) m->tt -= 1;
        }
        HIPCHK(hipGraphLaunch(e.tail, e.stream));
    }
    for (auto *m : e.members) m->tt += 1;
    return 0;
}

// cells of a group above which its steps run as streamed launches rather
// than as captured graphs (where a step's kernels take long enough that the
// host's launches keep ahead of the device); FREHD_GRAPHS=0 / 1 / 2 forces
// kernel-by-kernel steps / graphs / streamed launches
constexpr long long kStreamCells = 1 << 18;

int exec_step(Exec &e)
{
    if ((e.tim && e.tim->on) || e.count == 0) return exec_step_eager(e);
    int mode = e.useGraphs;
    if (mode == kLaunchAuto) {
        long long cells = 0;
        for (auto *m : e.members) cells += (long long)m->g.N;
        mode = cells >= kStreamCells ? kLaunchStreamed : kLaunchGraphs;
    }
    if (mode == kLaunchStreamed) return exec_step_streamed(e);
    // graphs once the group has stepped unchanged for a few steps (a group
    // whose members change every few steps would spend more on captures
    // than the graphs save); until then, and for a pattern's first step,
    // streamed launches (the same kernels, one wait for the flags)
    if (mode == kLaunchGraphs && ++e.stableSteps > 3) return exec_step_graphs(e);
    if (mode == kLaunchGraphs) return exec_step_streamed(e);
    return exec_step_eager(e);
}

// lets k_ssor<true> take `bytes` of dynamic LDS (beyond the default cap)
void ssor_allow_lds(int bytes)
{
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_ssor<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

int exec_init_subgrid(Exec &e)
{
    const Extents &x = e.ext;
    const int n = e.count;
    LAUNCH_ST(K_SUBINIT, k_subarea_init, grid1(x.nbN, n), dim3(256));
    LAUNCH_ST(K_COMBINE, k_combine, grid1(x.nbN + x.nbGhost, n), dim3(256));
    HIPCHK(hipStreamSynchronize(e.stream));
    return 0;
}

Extents member_extents(const frehd_dev_impl *d)
{
    const StepArgs &a = d->args;
    Extents x;
    x.gx2 = a.gx2;
    x.gy2 = a.gy2;
    x.blockX = d->blockX;
    x.nb1 = a.nb1;
    x.nbN = a.nbN;
    x.nbGhost = a.nbGhost;
    x.nbRes = a.nbRes;
    x.nbAsm = a.nbAsm;
    const bool mixed = d->cfg.precision == FREHD_PREC_MIXED;
    x.nbDir = mixed ? d->pcg32.nbDir : d->pcg.nbDir;
    x.nbRef = mixed ? d->ref.nb : 0;
    x.lanes = mixed ? d->pcg32.lanes : d->pcg.lanes;
    x.ssorN = d->pcg.ssor ? d->g.N : 0;
    x.warm = !mixed && d->pcg.warm;
    x.budget = d->args.p.budget;
    x.otf = d->cfg.useSubgrid == 1 && d->cfg.subgridOnTheFly == 1;
    x.ssorLds = !d->pcg.ssor || d->g.N <= d->pcg.ldsCells;
    return x;
}

static void merge(Exec &e, const frehd_dev_impl *d, bool first)
{
    const Extents m = member_extents(d);
    Extents &x = e.ext;
    if (first) {
        x = m;
        e.maxIter = d->cfg.maxIter;
        e.maxRefine = d->cfg.maxRefine;
        e.nScalarMax = d->args.p.nScalar;
        return;
    }
    x.gx2 = std::max(x.gx2, m.gx2);
    x.gy2 = std::max(x.gy2, m.gy2);
    x.nb1 = std::max(x.nb1, m.nb1);
    x.nbN = std::max(x.nbN, m.nbN);
    x.nbGhost = std::max(x.nbGhost, m.nbGhost);
    x.nbRes = std::max(x.nbRes, m.nbRes);
    x.nbAsm = std::max(x.nbAsm, m.nbAsm);
    x.nbDir = std::max(x.nbDir, m.nbDir);
    x.nbRef = std::max(x.nbRef, m.nbRef);
    x.ssorN = std::max(x.ssorN, m.ssorN);
    x.warm = x.warm || m.warm;
    x.budget = x.budget || m.budget;
    x.otf = x.otf || m.otf;
    x.ssorLds = x.ssorLds && m.ssorLds;
    e.maxIter = std::max(e.maxIter, d->cfg.maxIter);
    e.maxRefine = std::max(e.maxRefine, d->cfg.maxRefine);
    e.nScalarMax = std::max(e.nScalarMax, d->args.p.nScalar);
}

int fused_capacity(int device, int lanes)
{
    const void *k = lanes == 4 ? reinterpret_cast<const void *>(k_solve_mixed<4>)
                  : lanes == 2 ? reinterpret_cast<const void *>(k_solve_mixed<2>)
                               : reinterpret_cast<const void *>(k_solve_mixed<1>);
    int per = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, 256, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    return per * cus;
}

// The fused solve's plan for a mixed-precision group: each member gets
// blocks in proportion to its inner-solve tiles (one tile each when the
// device holds a block per tile of every member), within what the device
// holds at once (and fusedMax); a group with more members than blocks keeps
// the launched solve.
static int plan_fused(Exec &e)
{
    e.planBlocks = 0;
    if (!e.mixed || e.fusedMax < 0 || e.count == 0 || !e.dPlan || !e.dSync) return 0;
    int cap = std::min(fused_capacity(e.device, e.ext.lanes), e.planCap);
    if (e.fusedMax > 0) cap = std::min(cap, e.fusedMax);
    if (cap < e.count) return 0;
    long long total = 0;
    for (auto *d : e.members) total += d->pcg32.nbDir;
    e.hPlan.clear();
    for (int m = 0; m < e.count; m++) {
        const int nb = e.members[m]->pcg32.nbDir;
        long long B = total <= cap ? nb : std::max<long long>(1, (long long)nb * (cap - e.count) / total);
        B = std::min<long long>(B, nb);
        for (int j = 0; j < (int)B; j++) e.hPlan.push_back(SolveBlock{m, j, (int)B, 0});
    }
    if ((int)e.hPlan.size() > cap) return fail(FREHD_EHIP, "fused solve plan of %zu blocks exceeds %d",
                                               e.hPlan.size(), cap);
    HIPCHK(hipMemcpyAsync(e.dPlan, e.hPlan.data(), e.hPlan.size() * sizeof(SolveBlock), hipMemcpyHostToDevice,
                          e.stream));
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e.device) != hipSuccess || khz <= 0) khz = 100000;
    e.spinTicks = (unsigned long long)khz * 2000ull;   // a barrier waits 2 s at most
    e.planBlocks = (int)e.hPlan.size();
    return 0;
}

int exec_refresh(Exec &e)
{
    exec_release_graphs(e);
    e.count = (int)e.members.size();
    if (e.count > e.capacity) return fail(FREHD_EINVAL, "group holds %d members, capacity %d", e.count, e.capacity);
    for (int m = 0; m < e.count; m++) {
        const frehd_dev_impl *d = e.members[m];
        if (m == 0) {
            e.sub = d->cfg.useSubgrid == 1;
            e.mixed = d->cfg.precision == FREHD_PREC_MIXED;
            e.f32 = e.mixed;
        } else if ((d->cfg.useSubgrid == 1) != (bool)e.sub || (d->cfg.precision == FREHD_PREC_MIXED) != (bool)e.mixed) {
            return fail(FREHD_EINVAL, "a group's members must share useSubgrid and precision");
        }
        if (m > 0 && (d->pcg.ssor != 0) != (e.ext.ssorN > 0))
            return fail(FREHD_EINVAL, "a group's members must share the preconditioner");
        if (m > 0 && member_extents(d).lanes != e.ext.lanes)
            return fail(FREHD_EINVAL, "a group's members must share the solver's columns per lane");
        if (m > 0 && d->blockX != e.ext.blockX)
            return fail(FREHD_EINVAL, "a group's members must share the stencil kernels' block shape");
        merge(e, d, m == 0);
        HIPCHK(hipMemcpyAsync(e.dArgs + m, &d->args, sizeof(StepArgs), hipMemcpyHostToDevice, e.stream));
        HIPCHK(hipMemcpyAsync(e.dPcg + m, &d->pcg, sizeof(PcgArgs<double>), hipMemcpyHostToDevice, e.stream));
        HIPCHK(hipMemcpyAsync(e.dPcg32 + m, &d->pcg32, sizeof(PcgArgs<float>), hipMemcpyHostToDevice, e.stream));
        HIPCHK(hipMemcpyAsync(e.dRef + m, &d->ref, sizeof(RefineArgs), hipMemcpyHostToDevice, e.stream));
    }
    if (int rc = plan_fused(e)) return rc;
    HIPCHK(hipStreamSynchronize(e.stream));
    return 0;
}

int upload_member_args(frehd_dev_impl *d)
{
    d->self.members.assign(1, d);
    return exec_refresh(d->self);
}

}  // namespace frehd
