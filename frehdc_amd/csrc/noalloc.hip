// noalloc.hip -- frehd_dev_step_noalloc: whole time steps of a handle without
// a single allocation, for callers that cannot allocate (serving loops that
// must not call hipMalloc, code that forbids it).
//
// frehd_dev_step's kernels write their per-block partial sums (solver dot
// products, volume losses, CFL maxima) into buffers of the handle's arena.
// This variant points those kernels at one file-scope device array instead,
// `scratch`, of a fixed size, and runs the step's kernels one launch at a
// time (instantiating a captured graph allocates, so there are no graphs;
// there is no event timing either).  The handle's state and its other work
// arrays are the ones frehd_dev_step uses, so both entry points give the
// same results bit for bit and may be mixed on one handle.
//
// Every call on a device shares that device's copy of `scratch`: calls that
// overlap on one device (two host threads, two handles) race.  A call
// clears the part it uses before its first step.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "handle.h"

namespace frehd {

// partial sums of a handle's steps: 16 MiB holds those of grids up to about
// 4096 x 4096 cells (FP64 and mixed precision alike)
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

__global__ void __launch_bounds__(256) k_scratch_clear(double *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

namespace {

// each device's address of `scratch`, looked up once per device with that
// device current (the name is a host-side stand-in, not a device address)
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
char *g_scratchPtr[kMaxDevices];

int scratch_address(int device, char **out)
{
    if (device < 0 || device >= kMaxDevices) return fail(FREHD_EINVAL, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchPtr[device]) {
        void *p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return fail(FREHD_EHIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(err));
        g_scratchPtr[device] = static_cast<char *>(p);
    }
    *out = g_scratchPtr[device];
    return 0;
}

// the handle's argument blocks with their partial-sum pointers moved into
// scratch (starting at base); returns the bytes of scratch they take
size_t point_into_scratch(frehd_dev_impl *d, char *base)
{
    const PartialSizes ps = partial_sizes(d->cfg);
    size_t off = 0;
    auto take = [&](size_t n) {
        off = (off + 255) & ~size_t(255);
        // (typed as the argument blocks' pointers are in either compilation pass)
        gptr<double> p = reinterpret_cast<gptr<double>>(reinterpret_cast<uintptr_t>(base) + off);
        off += n * sizeof(double);
        return p;
    };
    StepArgs &a = d->naArgsH;
    a = d->args;
    a.partA = take(ps.partA);
    a.partB = take(ps.partB);
    // (the volume-loss accumulators and the CFL maxima the statistics read
    // stay in the handle's own memory)
    a.bpart = take(ps.bpart);
    PcgArgs<double> &p = d->naPcgH;
    p = d->pcg;
    p.partA = a.partA;
    p.partB = a.partB;
    if (d->cfg.precision != FREHD_PREC_MIXED && ps.partF) {
        p.partF[0] = take(ps.partF);
        p.partF[1] = take(ps.partF);
    }
    PcgArgs<float> &f = d->naPcg32H;
    f = d->pcg32;
    RefineArgs &r = d->naRefH;
    r = d->ref;
    if (d->cfg.precision == FREHD_PREC_MIXED) {
        f.partA = a.partA;
        f.partB = a.partB;
        f.partF[0] = take(ps.partF);
        f.partF[1] = take(ps.partF);
        r.partA = a.partA;
        r.partC = take(ps.partC);
    }
    return off;
}

}  // namespace
}  // namespace frehd

using namespace frehd;

extern "C" int frehd_dev_step_noalloc(frehd_dev_t *d, int32_t nsteps)
{
    if (!d) return fail(FREHD_EINVAL, "handle is NULL");
    if (nsteps < 0) return fail(FREHD_EINVAL, "nsteps < 0");
    if (step_call_check(d->midStep, d->batch != nullptr, CALL_STEP) != 0)
        return fail(FREHD_ESTATE, d->batch ? "the handle is in a batch: drive the batch"
                                           : "a step is half done: finish it with frehd_dev_phase");
    HIPCHK(hipSetDevice(d->device));
    char *base = nullptr;
    int rc = scratch_address(d->device, &base);
    if (rc) return rc;
    // size check on the host first (nothing written yet)
    const size_t need = point_into_scratch(d, nullptr);
    if (need > kScratchBytes)
        return fail(FREHD_EINVAL, "the step's partial sums need %zu bytes of scratch; frehd_dev_step_noalloc "
                                  "holds %zu (grids up to about 4096 x 4096 cells)", need, kScratchBytes);
    point_into_scratch(d, base);
    HIPCHK(hipStreamSynchronize(d->stream));
    Ctl c;
    HIPCHK(hipMemcpyAsync(&c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    if (c.halted || c.haltPending)
        return fail(FREHD_ECFL, "handle halted at step %d: max CFL exceeded 20", c.haltStep);
    if (c.tt + nsteps > d->cfg.Nt)
        return fail(FREHD_EINVAL, "stepping to %d exceeds the forcing series (Nt = %d)", c.tt + nsteps, d->cfg.Nt);
    if (nsteps == 0) return 0;
    HIPCHK(hipMemcpyAsync(d->naArgs, &d->naArgsH, sizeof(StepArgs), hipMemcpyHostToDevice, d->stream));
    HIPCHK(hipMemcpyAsync(d->naPcg, &d->naPcgH, sizeof(PcgArgs<double>), hipMemcpyHostToDevice, d->stream));
    HIPCHK(hipMemcpyAsync(d->naPcg32, &d->naPcg32H, sizeof(PcgArgs<float>), hipMemcpyHostToDevice, d->stream));
    HIPCHK(hipMemcpyAsync(d->naRef, &d->naRefH, sizeof(RefineArgs), hipMemcpyHostToDevice, d->stream));
    // the previous call (of any handle) left its sums here
    const size_t n = need / sizeof(double);
    const unsigned nb = (unsigned)std::min<size_t>(1024, (n + 255) / 256);
    hipLaunchKernelGGL(k_scratch_clear, dim3(nb), dim3(256), 0, d->stream, reinterpret_cast<double *>(base), n);
    HIPCHK(hipGetLastError());
    d->tt = c.tt;
    // the handle's group, launching with these argument blocks, one kernel
    // at a time, untimed; restored on the way out
    Exec &e = d->self;
    struct Restore {
        Exec &e;
        StepArgs *a;
        PcgArgs<double> *p;
        PcgArgs<float> *f;
        RefineArgs *r;
        int graphs;
        Timing *tim;
        ~Restore()
        {
            e.dArgs = a;
            e.dPcg = p;
            e.dPcg32 = f;
            e.dRef = r;
            e.useGraphs = graphs;
            e.tim = tim;
        }
    } restore{e, e.dArgs, e.dPcg, e.dPcg32, e.dRef, e.useGraphs, e.tim};
    e.dArgs = d->naArgs;
    e.dPcg = d->naPcg;
    e.dPcg32 = d->naPcg32;
    e.dRef = d->naRef;
    e.useGraphs = 0;
    e.tim = nullptr;
    for (int s = 0; s < nsteps; s++) {
        if ((rc = exec_step(e))) return rc;
        if (e.hFlags[F_HALTED]) break;
    }
    HIPCHK(hipStreamSynchronize(d->stream));
    HIPCHK(hipMemcpyAsync(&c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    if (c.halted || c.haltPending) {
        double mx = 0, my = 0;
        cfl_maxima(d, &mx, &my);
        return fail(FREHD_ECFL, "max CFL exceeded 20 at step %d (CFLx %.3g, CFLy %.3g)", c.haltStep, mx, my);
    }
    return 0;
}
