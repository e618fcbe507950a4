// api.hip -- C ABI of the device step (include/frehd_gpu.h): handle
// creation (Init restated on the host), device layout, upload, download,
// statistics, and groups of handles stepped together (frehd_batch_t).  The
// kernel sequencing itself is in exec.hip.
// host code only (no kernels): the argument blocks' pointers stay plain in
// both compilation passes
#define FREHD_HOST_SOURCE
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <string>
#include <vector>

#include "frehd_dev.h"
#include "handle.h"
#include "step_kernels.h"

using namespace frehd;

static thread_local std::string g_err;

namespace frehd {
int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
}  // namespace frehd

// ---------------------------------------------------------------------------
// Host restatement of Init (initialize.c:40-155) for a plain grid, np = 1.
// One-off setup on the host (as in the reference); the result is uploaded.
namespace {

struct HostState {
    int nx, ny, N, Ng, ns;
    std::vector<double> surf, surfOld, uuXP, vvYP, uuYP, vvXP, depth, depthXP, depthYP, EnXP,
        EnYP, DragXP, DragYP, CDXP, CDYP, Fuu, Fvv, cellV, S, Sm, CFLx, CFLy, wtfXP, wtfYP,
        bottomZ, bottomXP, bottomYP;
};

struct HG {
    Grid g;
    int E(int c) const { int i = c % g.nx, j = c / g.nx; return nbE(g, c, i, j); }
    int W(int c) const { int i = c % g.nx, j = c / g.nx; return nbW(g, c, i, j); }
    int Nn(int c) const { int i = c % g.nx, j = c / g.nx; return nbN(g, c, i, j); }
    int S(int c) const { int i = c % g.nx, j = c / g.nx; return nbS(g, c, i, j); }
};

void host_init(HostState &h, const frehd_config &k, const frehd_forcing &f, const frehd_ic &ic)
{
    const bool sub = k.useSubgrid == 1;
    const int nx = k.nx, ny = k.ny, N = nx * ny, Ng = (nx + 2) * (ny + 2);
    const int ns = std::max(1, k.nScalar);
    h.nx = nx; h.ny = ny; h.N = N; h.Ng = Ng; h.ns = ns;
    std::vector<double> *ng[] = {&h.surf, &h.surfOld, &h.uuXP, &h.vvYP, &h.uuYP, &h.vvXP,
        &h.depth, &h.depthXP, &h.depthYP, &h.EnXP, &h.EnYP, &h.DragXP, &h.DragYP, &h.CDXP,
        &h.CDYP, &h.Fuu, &h.Fvv, &h.cellV, &h.CFLx, &h.CFLy, &h.wtfXP, &h.wtfYP, &h.bottomZ,
        &h.bottomXP, &h.bottomYP};
    for (auto *v : ng) v->assign(Ng, 0.0);
    h.S.assign((size_t)ns * Ng, 0.0);
    h.Sm.assign((size_t)ns * Ng, 0.0);
    HG G{{nx, ny, N, Ng}};
    const Grid &g = G.g;
    double *zb = h.bottomZ.data(), *bx = h.bottomXP.data(), *by = h.bottomYP.data();
    // InitBathymetry + enforceBathBC + updateAllFaceElev (useCellEdge = 0)
    for (int c = 0; c < N; c++) zb[c] = ic.bottomZ[c] + k.offset;
    for (int i = 0; i < nx; i++) { zb[gjM(g, i)] = zb[i]; zb[gjP(g, i)] = zb[N - nx + i]; }
    for (int j = 0; j < ny; j++) { zb[giM(g, j)] = zb[j * nx]; zb[giP(g, j)] = zb[j * nx + nx - 1]; }
    // face bottoms: max of the two cells, or the subgrid edges (updateAllFaceElev
    // takes edgeX/edgeY when useSubgrid != 0, nsfunctions.c:84-96)
    for (int c = 0; c < N; c++) {
        if (sub) bx[c] = ic.sub->edgeX[c] + k.offset;
        else { double a = zb[c], b = zb[G.E(c)]; bx[c] = a > b ? a : b; }
    }
    for (int j = 0; j < ny; j++) { bx[giM(g, j)] = zb[j * nx]; bx[giP(g, j)] = bx[j * nx + nx - 1]; }
    for (int i = 0; i < nx; i++) { bx[gjM(g, i)] = bx[i]; bx[gjP(g, i)] = bx[N - nx + i]; }
    for (int c = 0; c < N; c++) {
        if (sub) by[c] = ic.sub->edgeY[c] + k.offset;
        else { double a = zb[c], b = zb[G.Nn(c)]; by[c] = a > b ? a : b; }
    }
    for (int i = 0; i < nx; i++) {
        double a = zb[gjM(g, i)], b = zb[i];
        by[gjM(g, i)] = a > b ? a : b;
        by[gjP(g, i)] = by[N - nx + i];
    }
    for (int j = 0; j < ny; j++) { by[giM(g, j)] = by[j * nx]; by[giP(g, j)] = by[j * nx + nx - 1]; }
    // initCD + initFieldValues
    for (int c = 0; c < Ng; c++) { h.CDXP[c] = k.CDx; h.CDYP[c] = k.CDy; }
    for (int c = 0; c < N; c++) {
        h.surf[c] = (ic.surf0cell ? ic.surf0cell[c] : ic.surf0) + k.offset;
        h.uuXP[c] = ic.U0cell ? ic.U0cell[c] : ic.U0;
        h.vvYP[c] = ic.V0cell ? ic.V0cell[c] : ic.V0;
        if (k.useScalar == 1)
            for (int q = 0; q < ns; q++)
                h.S[(size_t)q * Ng + c] = ic.S0cell ? ic.S0cell[(size_t)q * N + c] : ic.S0;
    }
    // zeroNegSurf(tt = 0)
    for (int c = 0; c < N; c++) if (h.surf[c] < zb[c]) h.surf[c] = zb[c];
    // enforceFreeSurfaceBC / enforceVelocityBC
    for (int kk = N; kk < Ng; kk++) {
        h.surf[kk] = h.surf[ghost_source(g, kk, false)];
        h.uuXP[kk] = h.uuXP[ghost_source(g, kk, true)];
        h.vvYP[kk] = h.vvYP[ghost_source(g, kk, false)];
    }
    // enforceTidalBC(tt = 0)
    if (k.bcType != 2)
        for (int q = 0; q < k.tideLocLengthP; q++) {
            int c = f.tideLocP[q];
            if (f.tideP[0] + k.offset > zb[c]) h.surf[c] = f.tideP[0] + k.offset;
        }
    if (k.bcType != 1)
        for (int q = 0; q < k.tideLocLengthM; q++) {
            int c = f.tideLocM[q];
            if (f.tideM[0] + k.offset > zb[c]) h.surf[c] = f.tideM[0] + k.offset;
        }
    // updateDepth + updateAllFaceDepth
    double *hs = h.surf.data(), *dX = h.depthXP.data(), *dY = h.depthYP.data();
    for (int c = 0; c < N; c++) { double d = hs[c] - zb[c]; h.depth[c] = d < 0 ? 0 : d; }
    for (int c = 0; c < N; c++) { double a = hs[c], b = hs[G.E(c)]; dX[c] = (a > b ? a : b) - bx[c]; }
    for (int j = 0; j < ny; j++) {
        dX[j * nx + nx - 1] = 0; dX[giP(g, j)] = 0; dX[giM(g, j)] = 0;
    }
    for (int i = 0; i < nx; i++) {
        int gm = gjM(g, i), gp = gjP(g, i);
        dX[gm] = (hs[gm] > hs[gm + 1] ? hs[gm] : hs[gm + 1]) - bx[gm];
        dX[gp] = (hs[gp] > hs[gp + 1] ? hs[gp] : hs[gp + 1]) - bx[gp];
    }
    for (int c = 0; c < N; c++) { double a = hs[c], b = hs[G.Nn(c)]; dY[c] = (a > b ? a : b) - by[c]; }
    for (int i = 0; i < nx; i++) { dY[gjP(g, i)] = hs[gjP(g, i)] - by[gjP(g, i)]; dY[gjM(g, i)] = 0; }
    for (int j = 0; j < ny; j++) { dY[giM(g, j)] = dY[j * nx]; dY[giP(g, j)] = dY[j * nx + nx - 1]; }
    for (int i = 0; i < nx; i++) { dY[N - nx + i] = 0; dY[gjP(g, i)] = 0; dY[gjM(g, i)] = 0; }
    for (int c = 0; c < Ng; c++) { if (dX[c] < 0) dX[c] = 0; if (dY[c] < 0) dY[c] = 0; }
    // computeVolume
    for (int c = 0; c < N; c++) {
        h.cellV[c] = k.dx * k.dy * h.depth[c];
        if (k.useScalar == 1)
            for (int q = 0; q < ns; q++) h.Sm[(size_t)q * Ng + c] = h.S[(size_t)q * Ng + c] * h.cellV[c];
    }
    // velocityInterp (flat-index neighbours; indices below 0 read 0)
    for (long c = 0; c < N; c++) {
        const double *u = h.uuXP.data(), *v = h.vvYP.data();
        double um1 = c - 1 >= 0 ? u[c - 1] : 0.0, um1n = c - 1 + nx >= 0 ? u[c - 1 + nx] : 0.0;
        double vmn = c - nx >= 0 ? v[c - nx] : 0.0, vmn1 = c + 1 - nx >= 0 ? v[c + 1 - nx] : 0.0;
        h.uuYP[c] = 0.25 * (u[c] + um1 + u[c + nx] + um1n);
        h.vvXP[c] = 0.25 * (v[c] + v[c + 1] + vmn + vmn1);
    }
    // thinLayerDrag (called unconditionally by Init)
    if (k.CDnotN == 1) {
        for (int c = 0; c < Ng; c++) {
            if (dX[c] <= k.hD) h.CDXP[c] = k.CDx + (k.CDmax - k.CDx) * (k.hD - dX[c]) / k.hD;
            if (dY[c] <= k.hD) h.CDYP[c] = k.CDy + (k.CDmax - k.CDy) * (k.hD - dY[c]) / k.hD;
        }
    } else {
        for (int c = 0; c < N; c++) {
            if (h.depth[c] > 0) {
                double n = f.manningCell ? f.manningCell[c] : k.manningN;
                double coeff = k.g * n * n;
                if (dX[c] <= k.hD) h.CDXP[c] = coeff / pow(h.depth[c], 2.0 / 3.0);
                if (dY[c] <= k.hD) h.CDYP[c] = coeff / pow(h.depth[c], 2.0 / 3.0);
            }
        }
    }
    // scalar: zero on dry cells, scalarMass
    if (k.useScalar == 1)
        for (int q = 0; q < ns; q++) {
            double *S = h.S.data() + (size_t)q * Ng, *Sm = h.Sm.data() + (size_t)q * Ng;
            for (int c = 0; c < N; c++) if (h.depth[c] <= 0) S[c] = 0;
            for (int c = 0; c < N; c++) Sm[c] = S[c] * h.cellV[c];
        }
}

}  // namespace

// Copies between host and device go on the handle's own stream and wait
// for it, never through the null stream: a synchronous null-stream copy from
// one host thread would interfere with another thread's stream capture.
static int copy_wait(hipStream_t st, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

#define COPYW(st, dst, src, bytes, kind)                        \
    do {                                                        \
        int rc_ = copy_wait((st), (dst), (src), (bytes), (kind)); \
        if (rc_) return rc_;                                    \
    } while (0)

namespace frehd {
int launch_subgrid_tables(const double *fineDev, void *tabs, int nx, int ny, int r, int Ne, double surfmin,
                          double dsurf, double dxf, int f32, hipStream_t stream);
int launch_tables_reorder(void *tabs, double *lm, int N, int Ne, int T, int toInterleaved, int f32, hipStream_t stream);
}

namespace {

struct Carve {
    char *base;
    size_t off;
    template <class T>
    T *take(size_t n)
    {
        off = (off + 255) & ~size_t(255);
        T *p = reinterpret_cast<T *>(base ? base + off : nullptr);
        off += n * sizeof(T);
        return p;
    }
};

// blocks of the grid-stride solver kernels: enough to fill the chip, few
// enough that every block can re-reduce all block partials cheaply
inline size_t pcg_blocks(size_t N) { return std::min<size_t>(1024, (N + 255) / 256); }

// Solver tiling: tiles of 256 W columns x H rows (W adjacent columns per
// lane: k_cg1, k_pcg1).  The default H gives about `target` tiles: 1024 for
// the SSOR path's k_pcg_dir (16 waves per CU on 256 CUs; one partial per
// tile for k_pcg_res to re-reduce), 512 for k_cg1 and k_pcg1, which write
// four or five partials per tile and re-reduce all of them in every tile of
// the next launch, a cost that grows with the square of the tile count.
// The XCD remap pads the tile count to a multiple of 8.
struct DirTiling {
    int gx, gy, H, remap, nb;
};

inline DirTiling dir_tiling(int nx, int ny, int H, int remap, int target = 1024, int W = 1)
{
    DirTiling t;
    t.gx = (nx + 256 * W - 1) / (256 * W);
    if (H <= 0) H = std::max(1, (int)(((int64_t)t.gx * ny + target - 1) / target));
    t.H = H;
    t.gy = (ny + H - 1) / H;
    t.remap = remap;
    t.nb = t.gx * t.gy;
    if (remap) t.nb = (t.nb + 7) & ~7;
    return t;
}

inline size_t max_dir_blocks(int nx, int ny) { return (size_t)((nx + 255) / 256) * ny + 8; }

// How a handle's or batch's whole steps are launched, from FREHD_GRAPHS in
// the environment at its creation: unset, streamed launches for groups of
// at least 2^18 cells and captured graphs below; 0 kernel by kernel with a
// wait for the solver flags (for profilers: under rocprofv3 a process that
// replayed the graphs of a second handle has died in hipGraphLaunch); 1
// graphs; 2 streamed launches
inline int graphs_from_env()
{
    const char *v = getenv("FREHD_GRAPHS");
    if (!v || !v[0]) return kLaunchAuto;
    if (v[0] == '0') return kLaunchEager;
    if (v[0] == '2') return kLaunchStreamed;
    return kLaunchGraphs;
}

// FREHD_PATTERN_CAP in the environment at a handle's or batch's creation:
// at most that many solver launches enqueued ahead (FP64 iterations, or the
// inner iterations of the mixed solve's first sweep, which is then its only
// pattern sweep), so that solves are left open and finished by the host --
// a test hook for that path; 0 or unset: the covering pattern
inline int pattern_cap_from_env()
{
    const char *v = getenv("FREHD_PATTERN_CAP");
    return v && v[0] ? std::max(0, atoi(v)) : 0;
}

// the most adjacent columns per lane the configuration's solver takes:
// k_cg1 4 (mixed), k_pcg1 2 (FP64 Jacobi), k_pcg_dir 1 (SSOR)
inline int max_lanes(const frehd_config &k)
{
    if (k.precision == FREHD_PREC_MIXED) return 4;
    return k.precond == FREHD_PRECOND_JACOBI ? 2 : 1;
}

// the group's flag array: pinned, coherent host memory mapped for the
// device, which the flag gather writes directly
static hipError_t alloc_flags(Exec &e, size_t members)
{
    const size_t bytes = members * F_STRIDE * sizeof(int32_t);
    hipError_t err = hipHostMalloc((void **)&e.hFlags, bytes, hipHostMallocMapped | hipHostMallocCoherent);
    if (err != hipSuccess) return err;
    memset(e.hFlags, 0, bytes);
    return hipHostGetDevicePointer((void **)&e.hFlagsDev, e.hFlags, 0);
}

// fine pixels per coarse cell side of the on-the-fly subgrid, 0 when subR
// is not a whole number
static int fine_r(const frehd_config &k)
{
    const double r = std::floor(k.subR + 0.5);
    if (!(r >= 1 && r <= 64 && std::fabs(k.subR - r) < 1e-9)) return 0;
    return (int)r;
}

// the combine reads minimum areas / volumes: useminA = 1, or useminV = 2
// (with useminA = 0 and useminV = 1 the reference's Nmin stays 0 and the
// volume adjustment never fires)
static bool min_tables(const frehd_config &k)
{
    return k.useSubgrid == 1 && (k.useminA == 1 || k.useminV == 2);
}

}  // namespace

namespace frehd {
PartialSizes partial_sizes(const frehd_config &k)
{
    const size_t N = (size_t)k.nx * k.ny, Ng = (size_t)(k.nx + 2) * (k.ny + 2);
    // 2-D sweeps: blocks of the largest grid any of the 256-thread shapes
    // (64 x 4, 128 x 2, 256 x 1) gives
    size_t nb2d = 0;
    for (int bx : {64, 128, 256})
        nb2d = std::max(nb2d, (size_t)((k.nx + bx - 1) / bx) * ((k.ny + 256 / bx - 1) / (256 / bx)));
    const size_t nb1d = (Ng + 255) / 256;
    const size_t G = pcg_blocks(N);
    const size_t maxb = std::max(std::max(nb2d, nb1d), G);
    const size_t maxDir = max_dir_blocks(k.nx, k.ny);
    PartialSizes z;
    z.maxBlocks = maxb;
    z.nb1 = nb1d;
    z.nb2 = nb2d;
    z.partA = 3 * std::max(G, maxDir);
    z.partB = maxDir;
    z.vpart = 4 * maxb;
    z.cflpart = 2 * maxb;
    const bool mixed = k.precision == FREHD_PREC_MIXED;
    z.partF = mixed ? 4 * maxDir : (k.precond == FREHD_PRECOND_JACOBI ? 5 * maxDir : 0);
    z.partC = mixed ? 2 * maxDir : 0;
    z.bpart = (size_t)budget_slots(k.useScalar == 1 ? std::max(1, k.nScalar) : 0) * maxb;
    return z;
}
}  // namespace frehd

namespace {

// the explicit state is stored in FP32 by the mixed form
inline bool state_f32(const frehd_config &k) { return k.precision == FREHD_PREC_MIXED; }

// lay out every device array; with base == nullptr only the size is computed
size_t layout(frehd_dev *d, char *base)
{
    const frehd_config &k = d->cfg;
    const size_t N = (size_t)k.nx * k.ny, Ng = (size_t)(k.nx + 2) * (k.ny + 2);
    const int ns = std::max(1, k.nScalar);
    const size_t Nt1 = (size_t)k.Nt + 1;
    Carve cv{base, 0};
    StepArgs &a = d->args;
    // the explicit state in the handle's state precision (FP32 for the mixed
    // form; the pointers are typed double on the host either way)
    const bool f32 = state_f32(k);
    auto takeR = [&](size_t n) { return f32 ? reinterpret_cast<double *>(cv.take<float>(n)) : cv.take<double>(n); };
    a.ctl = cv.take<Ctl>(1);
    double **ngArr[] = {&a.surfA, &a.surfB, &a.bottomZ, &a.bottomXP, &a.bottomYP};
    for (auto **p : ngArr) *p = cv.take<double>(Ng);
    double **ngR[] = {&a.uuXP, &a.vvYP, &a.depthXP, &a.depthYP, &a.EnXP, &a.EnYP, &a.DragXP, &a.DragYP,
        &a.CDXP, &a.CDYP, &a.Fuu, &a.Fvv, &a.CFLx, &a.CFLy, &a.wtfXP, &a.wtfYP};
    for (auto **p : ngR) *p = takeR(Ng);
    a.sStride = (int64_t)((Ng + 31) & ~size_t(31));
    a.S = takeR((size_t)a.sStride * ns);
    a.S2 = takeR((size_t)a.sStride * ns);
    a.Sm = cv.take<double>((size_t)a.sStride * ns);
    a.uuYP = takeR(N);
    a.vvXP = takeR(N);
    double **nArr[] = {&a.depthA, &a.depthB, &a.cellV, &a.cX, &a.cY, &a.D, &a.b, &a.x, &a.r, &a.q, &a.pA, &a.pB};
    for (auto **p : nArr) *p = cv.take<double>(N);
    // SSOR: the preconditioned residual z
    d->pcg.z = k.precond == FREHD_PRECOND_SSOR ? cv.take<double>(N) : nullptr;
    d->pcg.partF[0] = d->pcg.partF[1] = nullptr;
    a.manning = nullptr;
    a.rmv = cv.take<uint8_t>(N);
    a.flags = cv.take<uint8_t>(N);
    a.tface = cv.take<uint16_t>(N);
    const PartialSizes ps = partial_sizes(k);
    a.maxBlocks = (int)ps.maxBlocks;
    a.nb1 = (int)ps.nb1;
    a.nb2 = (int)ps.nb2;
    a.partA = cv.take<double>(ps.partA);
    a.partB = cv.take<double>(ps.partB);
    if (k.precision == FREHD_PREC_FP64 && ps.partF) {
        // k_pcg1's partials
        d->pcg.partF[0] = cv.take<double>(ps.partF);
        d->pcg.partF[1] = cv.take<double>(ps.partF);
    }
    a.cX32 = a.cY32 = nullptr;
    a.xB = nullptr;
    if (k.precision == FREHD_PREC_MIXED) {
        // scaled coefficients, two residual and two direction buffers, the
        // correction
        float **fArr[] = {&a.cX32, &a.cY32, &d->pcg32.r, &d->pcg32.q, &d->pcg32.pBuf[0],
                          &d->pcg32.pBuf[1], &d->pcg32.x};
        for (auto **p : fArr) *p = cv.take<float>(N);
        d->pcg32.partF[0] = cv.take<double>(ps.partF);
        d->pcg32.partF[1] = cv.take<double>(ps.partF);
        a.xB = cv.take<double>(N);
        d->ref.partC = cv.take<double>(ps.partC);
    }
    // the handle as a group of one: device copies of its argument blocks
    // and the flag gather buffer
    d->self.capacity = 1;
    d->self.dArgs = cv.take<StepArgs>(1);
    d->self.dPcg = cv.take<PcgArgs<double>>(1);
    d->self.dPcg32 = cv.take<PcgArgs<float>>(1);
    d->self.dRef = cv.take<RefineArgs>(1);
    // the fused mixed solve's plan and barrier state (zeroed with the arena)
    d->self.dSync = nullptr;
    d->self.dPlan = nullptr;
    d->self.planCap = 0;
    if (k.precision == FREHD_PREC_MIXED) {
        d->self.dSync = cv.take<SolveSync>(1);
        d->self.dPlan = cv.take<SolveBlock>(kPlanCap);
        d->self.planCap = kPlanCap;
    }
    // the argument blocks of frehd_dev_step_noalloc (partial sums in the
    // library's static scratch array instead of these)
    d->naArgs = cv.take<StepArgs>(1);
    d->naPcg = cv.take<PcgArgs<double>>(1);
    d->naPcg32 = cv.take<PcgArgs<float>>(1);
    d->naRef = cv.take<RefineArgs>(1);
    a.vpart = cv.take<double>(ps.vpart);
    a.cflpart = cv.take<double>(ps.cflpart);
    a.bpart = cv.take<double>(ps.bpart);
    a.tideP = cv.take<double>(Nt1);
    a.tideM = cv.take<double>(Nt1);
    a.inflow = cv.take<double>(Nt1);
    a.tidalPS = cv.take<double>(Nt1);
    a.tidalMS = cv.take<double>(Nt1);
    if (k.useWind == 1) {
        a.windSpd = cv.take<double>(Nt1);
        a.windCos = cv.take<double>(Nt1);
        a.windSin = cv.take<double>(Nt1);
    } else {
        a.windSpd = a.windCos = a.windSin = nullptr;
    }
    a.tideLocP = cv.take<int>(std::max(1, k.tideLocLengthP));
    a.tideLocM = cv.take<int>(std::max(1, k.tideLocLengthM));
    if (k.useSubgrid == 1) {
        double **ngs[] = {&a.V, &a.Z, &a.Vx, &a.Vy, &a.Zx, &a.Zy, &a.Nx, &a.Oy};
        for (auto **p : ngs) *p = takeR(Ng);
        double **ns2[] = {&a.Vxp, &a.Vxm, &a.Vyp, &a.Vym, &a.Np, &a.Nm, &a.Op, &a.Om, &a.Vraw, &a.Zraw};
        for (auto **p : ns2) *p = takeR(N);
        a.hSub = cv.take<double>(N);
        a.dSub = cv.take<uint8_t>(N);
        a.ind = cv.take<int>(Ng);
        a.edgeX = cv.take<double>(N);
        a.edgeY = cv.take<double>(N);
        a.allSurf = cv.take<double>((size_t)k.Ne);
        a.tabs = nullptr;
        a.bed = FineBed{nullptr, k.nx, k.ny, 0, 0.0};
        if (k.subgridOnTheFly == 1) {
            const int r = fine_r(k);
            a.bed.r = r;
            a.bed.z = cv.take<double>(N * (size_t)r * r);
        } else {
            a.tabs = takeR((size_t)k.Ne * N * 10);
        }
        a.allYhX = a.allYhY = a.allCn = nullptr;
        a.YhX = a.YhY = a.Cn = nullptr;
        const double **mt[8] = {&a.allNmin, &a.allNmax, &a.allOmin, &a.allOmax,
                                &a.allMinVxp, &a.allMinVxm, &a.allMinVyp, &a.allMinVym};
        double **mc[8] = {&a.Nmin, &a.Nmax, &a.Omin, &a.Omax, &a.minVxp, &a.minVxm, &a.minVyp, &a.minVym};
        for (int t = 0; t < 8; t++) { *mt[t] = nullptr; *mc[t] = nullptr; }
        if (min_tables(k))
            for (int t = 0; t < 8; t++) {
                if ((t < 4 && k.useminA == 1) || (t >= 4 && k.useminV != 0)) {
                    *mt[t] = cv.take<double>((size_t)k.Ne * N);
                    *mc[t] = cv.take<double>(N);
                }
            }
        if (k.useSubDrag == 2) {
            a.allYhX = cv.take<double>((size_t)k.Ne * N);
            a.allYhY = cv.take<double>((size_t)k.Ne * N);
            a.YhX = cv.take<double>(N);
            a.YhY = cv.take<double>(N);
        } else if (k.useSubDrag == 3 || k.useSubDrag == 4) {
            a.allCn = cv.take<double>((size_t)k.Ne * N);
            a.Cn = cv.take<double>(N);
        }
    }
    return cv.off + 256;
}

size_t layout_manning(frehd_dev *d, char *base, size_t off)
{
    Carve cv{base, off};
    double *m = cv.take<double>((size_t)d->cfg.nx * d->cfg.ny);
    if (base) d->args.manning = m;
    return cv.off + 256;
}

int check_config(const frehd_config *k)
{
    if (!k) return fail(FREHD_EINVAL, "config is NULL");
    if (k->nx < 4 || k->ny < 4) return fail(FREHD_EINVAL, "grid must be at least 4x4 (nx=%d ny=%d)", k->nx, k->ny);
    if ((int64_t)(k->nx + 2) * (k->ny + 2) >= (int64_t)1 << 31)
        return fail(FREHD_EINVAL, "grid too large for 32-bit cell indices");
    if (k->Nt < 1) return fail(FREHD_EINVAL, "Nt must be >= 1");
    if (!(k->dx > 0 && k->dy > 0 && k->dt > 0)) return fail(FREHD_EINVAL, "dx, dy, dt must be > 0");
    if (k->precond != FREHD_PRECOND_JACOBI && k->precond != FREHD_PRECOND_SSOR)
        return fail(FREHD_EINVAL, "unknown preconditioner %d", k->precond);
    if (k->precond == FREHD_PRECOND_SSOR && k->precision != FREHD_PREC_FP64)
        return fail(FREHD_EINVAL, "SSOR preconditions the FP64 solve only (the mixed path's inner CG is "
                                  "diagonally scaled)");
    if (k->precision != FREHD_PREC_FP64 && k->precision != FREHD_PREC_MIXED)
        return fail(FREHD_EINVAL, "unknown precision %d", k->precision);
    if (k->useWind != 0 && k->useWind != 1) return fail(FREHD_EINVAL, "useWind must be 0 or 1");
    if (k->useSubDrag != 0 && k->useSubDrag != 2 && k->useSubDrag != 3 && k->useSubDrag != 4)
        return fail(FREHD_EINVAL, "useSubDrag must be 0, 2, 3 or 4 (1 reads curvature tables the reference "
                                  "never fills)");
    if (k->useminA != 0 && k->useminA != 1) return fail(FREHD_EINVAL, "useminA must be 0 or 1");
    if (k->useminV < 0 || k->useminV > 2) return fail(FREHD_EINVAL, "useminV must be 0, 1 or 2");
    if (min_tables(*k) && k->subgridOnTheFly != 0)
        return fail(FREHD_EINVAL, "useminA = 1 / useminV = 2 need level tables (subgridOnTheFly = 0)");
    if (k->useSubDrag != 0 && (k->useSubgrid != 1 || k->subgridOnTheFly != 0))
        return fail(FREHD_EINVAL, "useSubDrag needs the subgrid with level tables (useSubgrid = 1, "
                                  "subgridOnTheFly = 0)");
    if (k->precision == FREHD_PREC_MIXED && (int64_t)k->nx * k->ny * 4 >= (int64_t)1 << 31)
        return fail(FREHD_EINVAL, "mixed precision addresses its FP32 vectors with 32-bit byte offsets "
                                  "(nx*ny < 2^29)");
    if (k->precision == FREHD_PREC_MIXED && !(k->epsInner > 0 && k->epsInner < 1 && k->maxRefine >= 1))
        return fail(FREHD_EINVAL, "mixed precision needs 0 < epsInner < 1 and maxRefine >= 1");
    if (k->warmStart < 0 || k->warmStart > 2) return fail(FREHD_EINVAL, "warmStart must be 0, 1 or 2");
    if (k->warmStart != 0 && k->precond == FREHD_PRECOND_SSOR)
        return fail(FREHD_EINVAL, "warmStart needs the Jacobi preconditioner (the SSOR path starts from 0, as the "
                                  "reference's solveMatrix does)");
    if (k->precision == FREHD_PREC_FP64 && k->precond == FREHD_PRECOND_JACOBI &&
        (int64_t)k->nx * k->ny * 8 > (int64_t)1 << 31)
        return fail(FREHD_EINVAL, "the FP64 solver addresses its vectors with 32-bit byte offsets (nx*ny <= 2^28)");
    if (k->useScalar == 1 && k->scalarAdv != 1 && k->scalarAdv != 2)
        return fail(FREHD_EINVAL, "scalarAdv must be 1 (first-order upwind) or 2 (TVD superbee)");
    if (k->bcType < 1 || k->bcType > 3) return fail(FREHD_EINVAL, "bcType must be 1, 2 or 3");
    if (k->tideLocLengthP < 0 || k->tideLocLengthM < 0 || k->inflowLocLength < 0)
        return fail(FREHD_EINVAL, "negative location count");
    if (k->useScalar == 1 && k->nScalar > FREHD_MAX_SCALARS)
        return fail(FREHD_EINVAL, "at most %d scalars", FREHD_MAX_SCALARS);
    if (k->useSubgrid != 0 && k->useSubgrid != 1) return fail(FREHD_EINVAL, "useSubgrid must be 0 or 1");
    if (k->useSubgrid == 1 && !(k->Ne >= 1 && k->dsurf > 0 && k->subR > 0))
        return fail(FREHD_EINVAL, "subgrid needs Ne >= 1, dsurf > 0, subR > 0");
    if (k->subgridOnTheFly != 0 && k->subgridOnTheFly != 1)
        return fail(FREHD_EINVAL, "subgridOnTheFly must be 0 or 1");
    if (k->subgridOnTheFly == 1 && k->useSubgrid != 1)
        return fail(FREHD_EINVAL, "subgridOnTheFly needs useSubgrid = 1");
    if (k->subgridOnTheFly == 1 && fine_r(*k) < 1)
        return fail(FREHD_EINVAL, "subgridOnTheFly needs an integral subR >= 1 (subR = %g)", k->subR);
    return 0;
}

int build_flags(const frehd_config &k, const frehd_forcing &f, std::vector<uint8_t> &flags)
{
    const int N = k.nx * k.ny;
    flags.assign(N, 0);
    for (int q = 0; q < k.tideLocLengthP; q++) {
        int c = f.tideLocP[q];
        if (c < 0 || c >= N) return fail(FREHD_EINVAL, "tideLocP[%d] = %d out of range", q, c);
        flags[c] |= 1;
    }
    for (int q = 0; q < k.tideLocLengthM; q++) {
        int c = f.tideLocM[q];
        if (c < 0 || c >= N) return fail(FREHD_EINVAL, "tideLocM[%d] = %d out of range", q, c);
        flags[c] |= 2;
    }
    for (int q = 0; q < k.inflowLocLength; q++) {
        int c = f.inflowLoc[q];
        if (c < 0 || c >= N) return fail(FREHD_EINVAL, "inflowLoc[%d] = %d out of range", q, c);
        if ((flags[c] >> 2) == 63) return fail(FREHD_EINVAL, "cell %d listed more than 63 times as inflow", c);
        flags[c] += 4;
    }
    return 0;
}

// Which vvYP faces adjustTidalVelocity (nsfunctions.c:1034-1078) sets, as
// the velocity interpolation of each cell c reads them, faces f = c, c + 1,
// c - nx, c + 1 - nx in bit pairs 2q, 2q + 1: the face's own cell is a
// tidal P cell (YP boundary), the cell north of it a tidal M cell (YM
// boundary); bit 8: c is a tidal M cell of the first row (its jM ghost face).
void build_tidal_faces(const frehd_config &k, const std::vector<uint8_t> &flags, std::vector<uint16_t> &tface)
{
    const int nx = k.nx, N = k.nx * k.ny;
    auto bits = [&](int f) -> int {
        if (f < 0 || f >= N) return 0;
        int b = 0;
        if (k.bcType != 2 && (flags[f] & 1)) b |= 1;
        if (k.bcType != 1 && f + nx < N && (flags[f + nx] & 2)) b |= 2;
        return b;
    };
    tface.assign(N, 0);
    for (int c = 0; c < N; c++) {
        int t = bits(c) | bits(c + 1) << 2 | bits(c - nx) << 4 | bits(c + 1 - nx) << 6;
        if (c < nx && k.bcType != 1 && (flags[c] & 2)) t |= 256;
        tface[c] = (uint16_t)t;
    }
}

void fill_params(P &p, const frehd_config &k)
{
    memset(&p, 0, sizeof p);
    p.dx = k.dx; p.dy = k.dy; p.dt = k.dt; p.g = k.g; p.NUx = k.NUx; p.NUy = k.NUy;
    p.CDx = k.CDx; p.CDy = k.CDy; p.manningN = k.manningN; p.hD = k.hD; p.CDmax = k.CDmax;
    p.minDepth = k.minDepth; p.wtfh = k.wtfh; p.CFLl = k.CFLl; p.CFLh = k.CFLh; p.eps = k.eps;
    p.tidePS = k.tidePS; p.tideMS = k.tideMS; p.inflowS = k.inflowS; p.Kx = k.Kx; p.Ky = k.Ky;
    p.offset = k.offset; p.CDnotN = k.CDnotN; p.useThinLayer = k.useThinLayer; p.bcType = k.bcType;
    p.tideLocLengthP = k.tideLocLengthP; p.tideLocLengthM = k.tideLocLengthM;
    p.inflowLocLength = k.inflowLocLength; p.useScalar = k.useScalar;
    p.nScalar = k.useScalar == 1 ? std::max(1, k.nScalar) : 0;
    p.useConstTidePS = k.useConstTidePS; p.useConstTideMS = k.useConstTideMS; p.maxIter = k.maxIter;
    p.scalarAdv = k.scalarAdv;
    p.useWind = k.useWind;
    p.useSubDrag = k.useSubgrid == 1 ? k.useSubDrag : 0;
    p.useminA = k.useminA;
    p.useminV = k.useminV;
    p.minTabs = min_tables(k) ? 1 : 0;
    p.beta = k.beta;
    p.rhoa = k.rhoa; p.Cw = k.Cw; p.CwT = k.CwT;
    p.useSubgrid = k.useSubgrid; p.dsurf = k.dsurf; p.subR = k.subR; p.Ne = k.Ne; p.surfmin = k.surfmin;
    p.warmStart = k.warmStart;
}

int upload(frehd_dev *d, const void *host, void *dev, size_t bytes)
{
    if (!bytes || !host) return 0;
    HIPCHK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, d->stream));
    return 0;
}

int upload_series(frehd_dev *d, const double *src, const double *dst)
{
    const size_t n = (size_t)d->cfg.Nt + 1;
    if (src) return upload(d, src, (void *)dst, n * sizeof(double));
    HIPCHK(hipMemsetAsync((void *)dst, 0, n * sizeof(double), d->stream));
    return 0;
}

void set_pcg_args(frehd_dev *d, int H, int remap, int lanes = 1);

// cells of k_ssor's vector that fit the device's LDS of one workgroup (a
// few hundred bytes stay for the kernel's own reductions)
int ssor_lds_cells(int device)
{
    int bytes = 0;
    if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        bytes = 65536;
    bytes = std::min(bytes, 160 * 1024) - 512;
    if (bytes <= 0) return 0;
    ssor_allow_lds(bytes);
    return bytes / (int)sizeof(double);
}

// this handle's launch extents (see StepArgs)
void fill_extents(frehd_dev *d)
{
    StepArgs &a = d->args;
    const frehd_config &k = d->cfg;
    const int bx = d->blockX, by = 256 / bx;
    a.gx2 = (k.nx + bx - 1) / bx;
    a.gy2 = (k.ny + by - 1) / by;
    a.nb2 = a.gx2 * a.gy2;
    a.nbN = (d->g.N + 255) / 256;
    a.nbGhost = (2 * k.nx + 2 * k.ny + 255) / 256;
    const int nt = (k.bcType != 2 ? k.tideLocLengthP : 0) + (k.bcType != 1 ? k.tideLocLengthM : 0);
    a.nbTide = (nt + 255) / 256;
    a.rowsDown = k.bcType != 2;
    a.nbRes = (int)pcg_blocks(d->g.N);
    // the assembly: one cell per thread, unless it also sums the FP64 cold
    // start's partials (r.r, r.z, b.b: nbRes of them, for the solver's first
    // launch to re-reduce)
    a.nbAsm = (k.precision == FREHD_PREC_MIXED || k.warmStart != 0) ? a.nbN : a.nbRes;
}

int create_common(const frehd_config *cfg, const frehd_forcing *f, frehd_dev **out, frehd_dev **dd)
{
    // (the handle gets a stream of its own, default priority)
    int rc = check_config(cfg);
    if (rc) return rc;
    if (!f || !out) return fail(FREHD_EINVAL, "forcing/out is NULL");
    if (cfg->bcType != 2 && cfg->tideLocLengthP > 0 && !(f->tideLocP && f->tideP))
        return fail(FREHD_EINVAL, "tideLocP/tideP missing");
    if (cfg->bcType != 1 && cfg->tideLocLengthM > 0 && !(f->tideLocM && f->tideM))
        return fail(FREHD_EINVAL, "tideLocM/tideM missing");
    if (cfg->inflowLocLength > 0 && !(f->inflowLoc && f->inflow))
        return fail(FREHD_EINVAL, "inflowLoc/inflow missing");
    if (cfg->useWind == 1 && !(f->windspd && f->winddir)) return fail(FREHD_EINVAL, "windspd/winddir missing");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FREHD_EHIP, "no HIP device visible");
    frehd_dev *d = new frehd_dev();
    d->cfg = *cfg;
    if (d->cfg.nScalar < 1) d->cfg.nScalar = 1;
    d->g = Grid{cfg->nx, cfg->ny, cfg->nx * cfg->ny, (cfg->nx + 2) * (cfg->ny + 2)};
    HIPCHK(hipGetDevice(&d->device));
    HIPCHK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    d->ownStream = true;
    size_t bytes = layout(d, nullptr);
    if (f->manningCell) bytes = layout_manning(d, nullptr, bytes);
    d->arenaBytes = bytes;
    if (hipMalloc(&d->arena, bytes) != hipSuccess) {
        hipStreamDestroy(d->stream);
        delete d;
        return fail(FREHD_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    HIPCHK(hipMemsetAsync(d->arena, 0, bytes, d->stream));
    size_t used = layout(d, d->arena);
    if (f->manningCell) layout_manning(d, d->arena, used);
    const int firstStep = 1;   // Ctl.ttNext at tt = 0
    COPYW(d->stream, &d->args.ctl->ttNext, &firstStep, sizeof(int), hipMemcpyHostToDevice);
    // the update phase may run first (a parity check entering a step there)
    const int go = 1;
    COPYW(d->stream, &d->args.ctl->runTail, &go, sizeof(int), hipMemcpyHostToDevice);
    // no surface seen yet by the subgrid lookup (all-ones bits: a NaN no
    // surface equals)
    if (cfg->useSubgrid == 1)
        HIPCHK(hipMemsetAsync(d->args.hSub, 0xFF, sizeof(double) * d->g.N, d->stream));
    d->args.g = d->g;
    fill_params(d->args.p, d->cfg);
    d->blockX = 64;
    fill_extents(d);
    set_pcg_args(d, 0, 0);
    d->self.device = d->device;
    d->self.stream = d->stream;
    d->self.tim = &d->timing;
    d->self.useGraphs = graphs_from_env();
    d->self.patternCap = pattern_cap_from_env();
    HIPCHK(alloc_flags(d->self, 1));
    d->midStep = 0;
    d->tt = 0;
    // forcing
    const StepArgs &a = d->args;
    if ((rc = upload_series(d, f->tideP, a.tideP))) return rc;
    if ((rc = upload_series(d, f->tideM, a.tideM))) return rc;
    if ((rc = upload_series(d, f->inflow, a.inflow))) return rc;
    if ((rc = upload_series(d, f->tidalPS, a.tidalPS))) return rc;
    if ((rc = upload_series(d, f->tidalMS, a.tidalMS))) return rc;
    if (cfg->useWind == 1) {
        // windTerm (nsfunctions.c:615-617): omega = (winddir + northAngle) pi / 180
        // with the reference's pi, and the host's cos / sin of it
        const size_t n = (size_t)cfg->Nt + 1;
        std::vector<double> cw(n), sw(n);
        for (size_t t = 0; t < n; t++) {
            const double pi = 3.1415926;
            const double phi = f->winddir[t] + cfg->northAngle;
            const double omega = phi * pi / 180.0;
            cw[t] = cos(omega);
            sw[t] = sin(omega);
        }
        if ((rc = upload_series(d, f->windspd, a.windSpd))) return rc;
        if ((rc = upload(d, cw.data(), (void *)a.windCos, n * sizeof(double)))) return rc;
        if ((rc = upload(d, sw.data(), (void *)a.windSin, n * sizeof(double)))) return rc;
        HIPCHK(hipStreamSynchronize(d->stream));   // the host vectors go out of scope
    }
    if (cfg->tideLocLengthP > 0 && f->tideLocP)
        if ((rc = upload(d, f->tideLocP, (void *)a.tideLocP, sizeof(int) * cfg->tideLocLengthP))) return rc;
    if (cfg->tideLocLengthM > 0 && f->tideLocM)
        if ((rc = upload(d, f->tideLocM, (void *)a.tideLocM, sizeof(int) * cfg->tideLocLengthM))) return rc;
    if (f->manningCell)
        if ((rc = upload(d, f->manningCell, d->args.manning, sizeof(double) * d->g.N))) return rc;
    std::vector<uint8_t> flags;
    if ((rc = build_flags(d->cfg, *f, flags))) return rc;
    if ((rc = upload(d, flags.data(), d->args.flags, flags.size()))) return rc;
    std::vector<uint16_t> tface;
    build_tidal_faces(d->cfg, flags, tface);
    if ((rc = upload(d, tface.data(), d->args.tface, tface.size() * sizeof(uint16_t)))) return rc;
    HIPCHK(hipStreamSynchronize(d->stream));
    if ((rc = upload_member_args(d))) return rc;
    *dd = d;
    return 0;
}


// n host doubles into a device array of the handle's state precision
// (rounded to FP32 for the mixed form; `keep` holds the rounded copies until
// the copies have completed)
int upload_r(frehd_dev *d, const double *src, double *dst, size_t n, std::vector<std::vector<float>> &keep)
{
    if (!src) return 0;
    if (!state_f32(d->cfg)) return upload(d, src, dst, n * sizeof(double));
    keep.emplace_back(n);
    std::vector<float> &f = keep.back();
    for (size_t i = 0; i < n; i++) f[i] = (float)src[i];
    return upload(d, f.data(), dst, n * sizeof(float));
}

// a device array of the handle's state precision into n host doubles
int download_r(frehd_dev *d, double *dst, const double *src, size_t n)
{
    if (!dst) return 0;
    if (!state_f32(d->cfg)) {
        COPYW(d->stream, dst, src, n * sizeof(double), hipMemcpyDeviceToHost);
        return 0;
    }
    std::vector<float> f(n);
    COPYW(d->stream, f.data(), src, n * sizeof(float), hipMemcpyDeviceToHost);
    for (size_t i = 0; i < n; i++) dst[i] = f[i];
    return 0;
}

// state arrays into the buffers the next step starts from (depthParity,
// sParity: the double-buffered depth and concentration buffers in use)
int upload_state(frehd_dev *d, const frehd_fields &s, int depthParity = 0, int sParity = 0)
{
    const size_t N = d->g.N, Ng = d->g.Ng;
    StepArgs &a = d->args;
    double *depthCur = depthParity ? a.depthB : a.depthA;
    std::vector<std::vector<float>> keep;
    // r: stored in the state precision
    struct { const double *src; double *dst; size_t n; bool r; } ng[] = {
        {s.surf, depthParity ? a.surfB : a.surfA, Ng, false}, {s.surfOld, depthParity ? a.surfA : a.surfB, Ng, false},
        {s.uuXP, a.uuXP, Ng, true},
        {s.vvYP, a.vvYP, Ng, true}, {s.uuYP, a.uuYP, N, true}, {s.vvXP, a.vvXP, N, true},
        {s.depth, depthCur, N, false}, {s.depthXP, a.depthXP, Ng, true}, {s.depthYP, a.depthYP, Ng, true},
        {s.EnXP, a.EnXP, Ng, true}, {s.EnYP, a.EnYP, Ng, true}, {s.DragXP, a.DragXP, Ng, true},
        {s.DragYP, a.DragYP, Ng, true}, {s.CDXP, a.CDXP, Ng, true}, {s.CDYP, a.CDYP, Ng, true},
        {s.Fuu, a.Fuu, Ng, true}, {s.Fvv, a.Fvv, Ng, true}, {s.cellV, a.cellV, N, false},
        {s.CFLx, a.CFLx, Ng, true}, {s.CFLy, a.CFLy, Ng, true}, {s.wtfXP, a.wtfXP, Ng, true},
        {s.wtfYP, a.wtfYP, Ng, true}, {s.bottomZ, a.bottomZ, Ng, false}, {s.bottomXP, a.bottomXP, Ng, false},
        {s.bottomYP, a.bottomYP, Ng, false}};
    for (auto &e : ng) {
        int rc = e.r ? upload_r(d, e.src, e.dst, e.n, keep) : upload(d, e.src, e.dst, e.n * sizeof(double));
        if (rc) return rc;
    }
    const int ns = d->args.p.nScalar;
    const size_t rw = state_f32(d->cfg) ? sizeof(float) : sizeof(double);
    for (int q = 0; q < ns; q++) {
        // both concentration buffers: the corners, which no step writes,
        // then agree whichever buffer is current (the current one last)
        if (s.S) {
            char *cur = reinterpret_cast<char *>(sParity ? a.S2 : a.S), *oth = reinterpret_cast<char *>(sParity ? a.S : a.S2);
            const size_t off = (size_t)q * a.sStride * rw;
            int rc = upload_r(d, s.S + (size_t)q * Ng, reinterpret_cast<double *>(oth + off), Ng, keep);
            if (!rc) rc = upload_r(d, s.S + (size_t)q * Ng, reinterpret_cast<double *>(cur + off), Ng, keep);
            if (rc) return rc;
        }
        if (s.Sm) { int rc = upload(d, s.Sm + (size_t)q * N, a.Sm + (size_t)q * a.sStride, N * sizeof(double)); if (rc) return rc; }
    }
    if (a.p.useSubgrid == 1) {
        struct { const double *src; double *dst; } sb[] = {{s.V, a.V}, {s.Z, a.Z}, {s.Vx, a.Vx},
            {s.Vy, a.Vy}, {s.Zx, a.Zx}, {s.Zy, a.Zy}, {s.Nx, a.Nx}, {s.Oy, a.Oy}};
        for (auto &e : sb) {
            int rc = upload_r(d, e.src, e.dst, Ng, keep);
            if (rc) return rc;
        }
        if (s.ind) { int rc = upload(d, s.ind, a.ind, Ng * sizeof(int)); if (rc) return rc; }
        // the lookup's skip cache holds results computed from the previous
        // bottoms and level index: forget every cell (all-ones bits, a NaN
        // no surface equals)
        HIPCHK(hipMemsetAsync(a.hSub, 0xFF, sizeof(double) * N, d->stream));
    }
    d->midStep = 0;
    HIPCHK(hipStreamSynchronize(d->stream));
    return 0;
}

// subgrid inputs: face bottoms (+ offset), level elevations, level tables
int upload_subgrid(frehd_dev *d, const frehd_subgrid *sg)
{
    const frehd_config &k = d->cfg;
    StepArgs &a = d->args;
    const size_t N = d->g.N, tn = (size_t)k.Ne * N;
    if (!sg || !sg->edgeX || !sg->edgeY) return fail(FREHD_EINVAL, "subgrid tables missing");
    const double *tabs[10] = {sg->V, sg->Z, sg->Vxp, sg->Vxm, sg->Vyp, sg->Vym, sg->Np, sg->Nm, sg->Op, sg->Om};
    const bool onTheFly = k.subgridOnTheFly == 1;
    const bool generate = sg->fine != nullptr && (!sg->V || onTheFly);
    if (generate && !(sg->fineR >= 1 && sg->dxf > 0))
        return fail(FREHD_EINVAL, "fine bathymetry needs fineR >= 1 and dxf > 0");
    if (onTheFly && !(generate && sg->fineR == a.bed.r))
        return fail(FREHD_EINVAL, "subgridOnTheFly needs the fine bathymetry with fineR = subR");
    for (int t = 0; t < 10 && !generate; t++)
        if (!tabs[t]) return fail(FREHD_EINVAL, "subgrid table %d missing", t);
    std::vector<double> ex(N), ey(N), lev(k.Ne);
    for (size_t c = 0; c < N; c++) { ex[c] = sg->edgeX[c] + k.offset; ey[c] = sg->edgeY[c] + k.offset; }
    for (int q = 0; q < k.Ne; q++) lev[q] = k.surfmin + q * k.dsurf + k.offset;
    int rc;
    if ((rc = upload(d, ex.data(), (void *)a.edgeX, N * sizeof(double)))) return rc;
    if ((rc = upload(d, ey.data(), (void *)a.edgeY, N * sizeof(double)))) return rc;
    if ((rc = upload(d, lev.data(), (void *)a.allSurf, k.Ne * sizeof(double)))) return rc;
    if (k.useSubDrag == 2) {
        if (!sg->YhX || !sg->YhY) return fail(FREHD_EINVAL, "useSubDrag = 2 needs the YhX and YhY tables");
        if ((rc = upload(d, sg->YhX, (void *)a.allYhX, tn * sizeof(double)))) return rc;
        if ((rc = upload(d, sg->YhY, (void *)a.allYhY, tn * sizeof(double)))) return rc;
    } else if (k.useSubDrag == 3 || k.useSubDrag == 4) {
        if (!sg->Cn) return fail(FREHD_EINVAL, "useSubDrag = %d needs the Cn table", k.useSubDrag);
        if ((rc = upload(d, sg->Cn, (void *)a.allCn, tn * sizeof(double)))) return rc;
    }
    if (min_tables(k)) {
        const double *src[8] = {sg->Nmin, sg->Nmax, sg->Omin, sg->Omax, sg->minVxp, sg->minVxm, sg->minVyp, sg->minVym};
        const double *dst[8] = {a.allNmin, a.allNmax, a.allOmin, a.allOmax,
                                a.allMinVxp, a.allMinVxm, a.allMinVyp, a.allMinVym};
        for (int t = 0; t < 8; t++) {
            if (!dst[t]) continue;
            if (!src[t]) return fail(FREHD_EINVAL, "useminA / useminV need the %s table",
                                     t < 4 ? "Nmin/Nmax/Omin/Omax" : "minV*");
            if ((rc = upload(d, src[t], (void *)dst[t], tn * sizeof(double)))) return rc;
        }
    }
    if (onTheFly) {
        // the argument block went to the device before the pixel size was
        // known: send it again
        a.bed.dxf = sg->dxf;
        rc = upload(d, sg->fine, (void *)a.bed.z, N * (size_t)a.bed.r * a.bed.r * sizeof(double));
        return rc ? rc : upload_member_args(d);
    }
    if (generate) {
        const size_t nf = N * (size_t)sg->fineR * sg->fineR;
        double *fine = nullptr;
        if (hipMalloc(&fine, nf * sizeof(double)) != hipSuccess)
            return fail(FREHD_ENOMEM, "hipMalloc of the fine bathymetry (%zu bytes) failed", nf * sizeof(double));
        rc = upload(d, sg->fine, fine, nf * sizeof(double));
        if (!rc && launch_subgrid_tables(fine, const_cast<double *>(a.tabs), k.nx, k.ny, sg->fineR, k.Ne,
                                         k.surfmin, k.dsurf, sg->dxf, state_f32(k), d->stream))
            rc = fail(FREHD_EHIP, "table generator launch failed");
        hipStreamSynchronize(d->stream);
        hipFree(fine);
        return rc;
    }
    // level-major host tables -> staging -> interleaved
    double *stage = nullptr;
    if (hipMalloc(&stage, tn * sizeof(double)) != hipSuccess)
        return fail(FREHD_ENOMEM, "hipMalloc of the table staging buffer (%zu bytes) failed", tn * sizeof(double));
    for (int t = 0; t < 10 && !rc; t++) {
        rc = upload(d, tabs[t], stage, tn * sizeof(double));
        if (!rc && launch_tables_reorder(const_cast<double *>(a.tabs), stage, (int)N, k.Ne, t, 1, state_f32(k),
                                         d->stream))
            rc = fail(FREHD_EHIP, "table reorder launch failed");
    }
    hipStreamSynchronize(d->stream);
    hipFree(stage);
    return rc;
}

void set_pcg_args(frehd_dev *d, int H, int remap, int lanes)
{
    d->tileH = H;
    d->tileRemap = remap;
    d->tileLanes = lanes;
    const StepArgs &a = d->args;
    PcgArgs<double> &p = d->pcg;
    p.cX = a.cX;
    p.cY = a.cY;
    p.D = a.D;
    p.r = a.r;
    p.q = a.q;
    p.pBuf[0] = a.pA;
    p.pBuf[1] = a.pB;
    p.x = a.x;
    p.partA = a.partA;
    p.partB = a.partB;
    p.sc = &a.ctl->pcg;
    p.halted = &a.ctl->halted;
    p.eps = d->cfg.eps;
    p.maxIter = d->cfg.maxIter;
    p.nx = d->g.nx;
    p.ny = d->g.ny;
    p.N = d->g.N;
    // FP64: k_pcg1 (Jacobi) takes W columns per lane and writes five
    // partials per tile, k_pcg_dir (SSOR) one column and one partial
    const bool mixed = d->cfg.precision == FREHD_PREC_MIXED;
    const bool jacobi64 = !mixed && d->cfg.precond == FREHD_PRECOND_JACOBI;
    p.lanes = jacobi64 ? lanes : 1;
    const DirTiling t = dir_tiling(p.nx, p.ny, H, remap, d->cfg.precond == FREHD_PRECOND_SSOR ? 1024 : 512, p.lanes);
    p.gx = t.gx;
    p.gy = t.gy;
    p.H = t.H;
    p.remap = t.remap;
    p.nbDir = t.nb;
    p.b = a.b;
    p.warm = !mixed && d->cfg.warmStart != 0;
    p.nbRes = (int)pcg_blocks(p.N);
    p.nbStart = p.nbRes;
    p.outerDone = nullptr;
    p.epsDyn = nullptr;
    p.ssor = d->cfg.precond == FREHD_PRECOND_SSOR ? 1 : 0;
    p.ldsCells = ssor_lds_cells(d->device);
    if (d->cfg.precision != FREHD_PREC_MIXED) return;
    // inner FP32 solve: same tiling, FP32 operator and vectors (r, q, p, x =
    // the correction d were carved by layout()), started by k_refine
    PcgArgs<float> &f = d->pcg32;
    f.cX = a.cX32;
    f.cY = a.cY32;
    f.D = nullptr;
    f.partA = a.partA;
    f.partB = a.partB;
    f.sc = &a.ctl->pcgIn;
    f.halted = &a.ctl->halted;
    f.outerDone = &a.ctl->ref.doneAt;
    f.epsDyn = &a.ctl->ref.epsIn;
    f.eps = d->cfg.epsInner;
    f.maxIter = p.maxIter;
    f.nx = p.nx; f.ny = p.ny; f.N = p.N;
    const DirTiling tc = dir_tiling(p.nx, p.ny, H, remap, 512, lanes);
    f.gx = tc.gx; f.gy = tc.gy; f.H = tc.H; f.remap = tc.remap;
    f.lanes = lanes;
    f.nbDir = tc.nb;
    f.nbRes = p.nbRes;
    // the refinement walks its own tiles (about 1024: it has the bytes of a
    // stencil sweep and four partials per tile, re-reduced by one block and
    // by the inner solve's set-up launch only); the set-up launch reduces
    // its partials
    const DirTiling tr = dir_tiling(p.nx, p.ny, 0, remap, 1024);
    f.nbStart = tr.nb;
    RefineArgs &r = d->ref;
    r.cX = a.cX; r.cY = a.cY; r.D = a.D; r.b = a.b;
    r.xbuf[0] = a.x;
    r.xbuf[1] = a.xB;
    r.d = f.x;
    r.r32 = f.r;
    r.cXs = a.cX32;
    r.cYs = a.cY32;
    r.partA = a.partA;
    r.ctl = a.ctl;
    r.eps = d->cfg.eps;
    r.epsInner = d->cfg.epsInner;
    r.maxRefine = d->cfg.maxRefine;
    r.nx = p.nx; r.ny = p.ny; r.N = p.N;
    r.gx = tr.gx; r.gy = tr.gy; r.H = tr.H; r.remap = tr.remap;
    r.nb = tr.nb;
}

// Init's subgrid part (initialize.c:103-124): level index at round(Ne/2)
// (integer division first, subgrid.c:2272), computeSubArea, combineSubArea
int init_subgrid_device(frehd_dev *d)
{
    std::vector<int> ind(d->g.Ng, d->cfg.Ne / 2);
    int rc = upload(d, ind.data(), d->args.ind, ind.size() * sizeof(int));
    if (rc) return rc;
    return exec_init_subgrid(d->self);
}

// Solver tiling of a group chosen by timing: every candidate (rows per tile
// 1, 2, 4, ..., up to the smallest member height, and the default; XCD remap
// off and on; more columns per lane where the solver and every member's nx
// allow it: 2 and 4 for the mixed path's inner solve, 2 for the FP64 Jacobi
// solve) runs `iters` solver iterations on the members' current systems with
// a zero tolerance, twice, and the fastest (by its better time) is applied
// to every member.
// Control blocks (statistics, step counter, flags) are restored, and the
// solver's work arrays are rewritten at the next step's assembly, so the
// model state is untouched.
int autotune_group(Exec &e, int iters, int32_t *bestH, int32_t *bestRemap, int32_t *bestLanes, double *bestMs)
{
    if (e.count == 0) return fail(FREHD_ESTATE, "nothing to tune");
    if (iters < 1 || iters > 1000) return fail(FREHD_EINVAL, "iters %d outside [1, 1000]", iters);
    HIPCHK(hipSetDevice(e.device));
    HIPCHK(hipStreamSynchronize(e.stream));
    std::vector<Ctl> saved(e.count);
    int nyMin = 1 << 30;
    for (int m = 0; m < e.count; m++) {
        COPYW(e.stream, &saved[m], e.members[m]->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
        nyMin = std::min(nyMin, e.members[m]->g.ny);
    }
    std::vector<int> Hs = {0};
    for (int H = 1; H <= nyMin && H <= 256; H *= 2) Hs.push_back(H);
    std::vector<int> Ws = {1};
    for (int W : {2, 4}) {
        bool ok = true;
        for (auto *d : e.members)
            if (d->g.nx % W || W > max_lanes(d->cfg)) ok = false;
        if (ok) Ws.push_back(W);
    }
    // every candidate twice, in two passes, its faster time counting (a
    // single short timing is at the mercy of clock and cache noise)
    struct Cand {
        int H, remap, W;
        double ms;
    };
    std::vector<Cand> cands;
    for (int W : Ws)
        for (int H : Hs)
            for (int remap = 0; remap < 2; remap++) {
                bool ok = true;
                for (auto *d : e.members)
                    if ((size_t)dir_tiling(d->g.nx, d->g.ny, H, remap).nb > max_dir_blocks(d->g.nx, d->g.ny))
                        ok = false;
                if (ok) cands.push_back(Cand{H, remap, W, -1.0});
            }
    int rc = 0, bH = 0, bR = 0, bW = 1;
    double best = -1.0;
    for (int pass = 0; pass < 2 && !rc; pass++)
        for (Cand &k : cands) {
            for (auto *d : e.members) {
                set_pcg_args(static_cast<frehd_dev *>(d), k.H, k.remap, k.W);
                d->pcg.eps = 0.0;                     // the trial iterations must all run
                d->pcg.maxIter = 1 << 30;
                d->pcg32.maxIter = 1 << 30;
            }
            if ((rc = exec_refresh(e))) break;
            for (int m = 0; m < e.count && !rc; m++) {
                Ctl c = saved[m];
                c.halted = 0;
                c.pcg.doneAt = -1;
                c.pcgIn.doneAt = -1;
                c.ref.doneAt = -1;
                c.ref.epsIn = 0.0;
                rc = copy_wait(e.stream, e.members[m]->args.ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice);
            }
            double ms = 0.0;
            if (!rc) rc = exec_time_solver(e, iters, &ms);
            if (rc) break;
            if (ms > 0 && (k.ms < 0 || ms < k.ms)) k.ms = ms;
        }
    for (const Cand &k : cands)
        if (k.ms > 0 && (best < 0 || k.ms < best)) {
            best = k.ms;
            bH = k.H;
            bR = k.remap;
            bW = k.W;
        }
    // the winner (or, after an error, the default), then the saved control blocks
    for (auto *d : e.members) set_pcg_args(static_cast<frehd_dev *>(d), rc ? 0 : bH, rc ? 0 : bR, rc ? 1 : bW);
    int rc2 = exec_refresh(e);
    for (int m = 0; m < e.count && !rc2; m++)
        rc2 = copy_wait(e.stream, e.members[m]->args.ctl, &saved[m], sizeof(Ctl), hipMemcpyHostToDevice);
    if (rc) return rc;
    if (rc2) return rc2;
    if (bestH) *bestH = e.mixed ? e.members[0]->pcg32.H : e.members[0]->pcg.H;
    if (bestRemap) *bestRemap = e.members[0]->pcg.remap;
    if (bestLanes) *bestLanes = e.mixed ? e.members[0]->pcg32.lanes : e.members[0]->pcg.lanes;
    if (bestMs) *bestMs = best;
    return 0;
}

}  // namespace

// FREHD_ESTATE (with a message) when `call` is not valid in the handle's
// current place inside a step or while it belongs to a batch
static int state_check(const frehd_dev *d, StepCall call, int phase = 0)
{
    const int rc = step_call_check(d->midStep, d->batch != nullptr, call, phase);
    if (rc == FREHD_EINVAL) return fail(rc, "phase %d out of range [0, 4]", phase);
    if (rc == 0) return 0;
    if (d->batch) return fail(rc, "the handle is in a batch: drive the batch");
    if (call == CALL_PHASE)
        return fail(rc, "phase %d again: %d phase(s) of the current step have run", phase, d->midStep);
    return fail(rc, "a step is half done (%d of its 5 phases ran): finish it with frehd_dev_phase", d->midStep);
}

// ---------------------------------------------------------------------------
extern "C" {

const char *frehd_last_error(void) { return g_err.c_str(); }

const char *frehd_version(void) { return "frehd-mi355x 0.3 (gfx950, plain and subgrid step, Jacobi PCG in FP64 or FP32 with FP64 refinement)"; }

int frehd_device_count(int32_t *count)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
    return 0;
}

int frehd_abi_sizes(int64_t sizes[6])
{
    sizes[0] = sizeof(frehd_config);
    sizes[1] = sizeof(frehd_forcing);
    sizes[2] = sizeof(frehd_ic);
    sizes[3] = sizeof(frehd_fields);
    sizes[4] = sizeof(frehd_stats);
    sizes[5] = sizeof(frehd_budget);
    return 0;
}

int frehd_dev_footprint(const frehd_config *cfg, int64_t *bytes)
{
    int rc = check_config(cfg);
    if (rc) return rc;
    frehd_dev tmp;
    tmp.cfg = *cfg;
    if (tmp.cfg.nScalar < 1) tmp.cfg.nScalar = 1;
    size_t b = layout(&tmp, nullptr);
    b = layout_manning(&tmp, nullptr, b);
    *bytes = (int64_t)b;
    return 0;
}

int frehd_dev_create(const frehd_config *cfg, const frehd_forcing *f, const frehd_ic *ic, frehd_dev_t **out)
{
    if (!ic || !ic->bottomZ) return fail(FREHD_EINVAL, "ic/bottomZ is NULL");
    if (cfg && f) {
        if (cfg->bcType != 2 && cfg->tideLocLengthP > 0 && !(f->tideLocP && f->tideP))
            return fail(FREHD_EINVAL, "tideLocP/tideP missing");
        if (cfg->bcType != 1 && cfg->tideLocLengthM > 0 && !(f->tideLocM && f->tideM))
            return fail(FREHD_EINVAL, "tideLocM/tideM missing");
    }
    int rc = check_config(cfg);
    if (rc) return rc;
    if (!f) return fail(FREHD_EINVAL, "forcing is NULL");
    if (cfg->useSubgrid == 1 && !(ic->sub && ic->sub->edgeX && ic->sub->edgeY))
        return fail(FREHD_EINVAL, "subgrid inputs (ic->sub) missing");
    HostState h;
    host_init(h, *cfg, *f, *ic);
    frehd_dev *d = nullptr;
    if ((rc = create_common(cfg, f, out, &d))) return rc;
    frehd_fields s;
    memset(&s, 0, sizeof s);
    s.surf = h.surf.data(); s.surfOld = h.surfOld.data(); s.uuXP = h.uuXP.data(); s.vvYP = h.vvYP.data();
    s.uuYP = h.uuYP.data(); s.vvXP = h.vvXP.data(); s.depth = h.depth.data(); s.depthXP = h.depthXP.data();
    s.depthYP = h.depthYP.data(); s.EnXP = h.EnXP.data(); s.EnYP = h.EnYP.data(); s.DragXP = h.DragXP.data();
    s.DragYP = h.DragYP.data(); s.CDXP = h.CDXP.data(); s.CDYP = h.CDYP.data(); s.Fuu = h.Fuu.data();
    s.Fvv = h.Fvv.data(); s.cellV = h.cellV.data(); s.CFLx = h.CFLx.data(); s.CFLy = h.CFLy.data();
    s.wtfXP = h.wtfXP.data(); s.wtfYP = h.wtfYP.data(); s.bottomZ = h.bottomZ.data();
    s.bottomXP = h.bottomXP.data(); s.bottomYP = h.bottomYP.data();
    // S in Ng stride, Sm in N stride for upload_state
    std::vector<double> Sm((size_t)h.ns * h.N);
    for (int q = 0; q < h.ns; q++)
        memcpy(Sm.data() + (size_t)q * h.N, h.Sm.data() + (size_t)q * h.Ng, sizeof(double) * h.N);
    s.S = h.S.data();
    s.Sm = Sm.data();
    if ((rc = upload_state(d, s))) { frehd_dev_destroy(d); return rc; }
    if (cfg->useSubgrid == 1) {
        if ((rc = upload_subgrid(d, ic->sub)) || (rc = init_subgrid_device(d))) {
            frehd_dev_destroy(d);
            return rc;
        }
    }
    *out = d;
    return 0;
}

int frehd_dev_create_from_state(const frehd_config *cfg, const frehd_forcing *f, const frehd_subgrid *sub,
                                const frehd_fields *state, frehd_dev_t **out)
{
    if (!state) return fail(FREHD_EINVAL, "state is NULL");
    if (cfg && cfg->useSubgrid == 1) {
        const double *sreq[] = {state->V, state->Z, state->Vx, state->Vy, state->Zx, state->Zy,
                                state->Nx, state->Oy};
        for (auto *p : sreq) if (!p) return fail(FREHD_EINVAL, "subgrid state array missing");
        if (!state->ind) return fail(FREHD_EINVAL, "subgrid level index missing");
    }
    const double *req[] = {state->surf, state->surfOld, state->uuXP, state->vvYP, state->uuYP, state->vvXP,
        state->depth, state->depthXP, state->depthYP, state->CDXP, state->CDYP, state->cellV,
        state->CFLx, state->CFLy, state->bottomZ, state->bottomXP, state->bottomYP};
    for (auto *p : req) if (!p) return fail(FREHD_EINVAL, "state array missing");
    if (cfg && cfg->useScalar == 1 && !(state->S && state->Sm)) return fail(FREHD_EINVAL, "state S/Sm missing");
    frehd_dev *d = nullptr;
    int rc = create_common(cfg, f, out, &d);
    if (rc) return rc;
    if ((rc = upload_state(d, *state))) { frehd_dev_destroy(d); return rc; }
    if (cfg->useSubgrid == 1 && (rc = upload_subgrid(d, sub))) { frehd_dev_destroy(d); return rc; }
    *out = d;
    return 0;
}

}  // extern "C"

namespace frehd {

// the maxima of the last step's CFL numbers, from the velocity update's
// block partials
int cfl_maxima(frehd_dev_impl *d, double *mx, double *my)
{
    const StepArgs &a = d->args;
    const size_t nb = (size_t)a.nb1;
    std::vector<double> v(nb);
    double m[2] = {0.0, 0.0};
    for (int q = 0; q < 2; q++) {
        COPYW(d->stream, v.data(), a.cflpart + (size_t)q * a.maxBlocks, nb * sizeof(double), hipMemcpyDeviceToHost);
        for (double x : v) m[q] = std::fmax(m[q], x);
    }
    *mx = m[0];
    *my = m[1];
    return 0;
}

// limiterCFL's / zeroNegSurf's / the velocity update's volume losses since
// creation: the per-block accumulators summed in block order
static int vloss_totals(frehd_dev_impl *d, double out[4])
{
    const StepArgs &a = d->args;
    std::vector<double> v((size_t)a.maxBlocks);
    for (int q = 0; q < 4; q++) {
        COPYW(d->stream, v.data(), a.vpart + (size_t)q * a.maxBlocks, v.size() * sizeof(double), hipMemcpyDeviceToHost);
        double t = 0.0;
        for (double x : v) t += x;
        out[q] = t;
    }
    return 0;
}

// the buffers holding the current depth, concentration and surface (1:
// depthB / S2 / surfB, whose partner then holds surfOld): between steps the
// device's start-of-step buffers with the flips the last step left pending;
// between the phases of a step that is not halted, the new ones once the
// solve (surface), updateData (depth) or scalarTransport (S) has run
static void current_buffers(const frehd_dev_impl *d, const Ctl &c, int *dp, int *sp, int *hp)
{
    const bool scal = d->args.p.nScalar > 0;
    if (d->midStep != 0 && !c.halted) {
        *dp = c.parity ^ (step_state_new_depth(d->midStep) ? 1 : 0);
        *sp = c.sParity ^ (step_state_new_scalar(d->midStep) ? 1 : 0);
        *hp = c.parity ^ (step_state_new_surface(d->midStep) ? 1 : 0);
    } else {
        *dp = c.parity ^ (c.flipPending ? 1 : 0);
        *sp = c.sParity ^ (c.flipPending && scal ? 1 : 0);
        *hp = *dp;
    }
}

}  // namespace frehd

extern "C" {

int frehd_dev_step(frehd_dev_t *d, int32_t nsteps)
{
    if (!d) return fail(FREHD_EINVAL, "handle is NULL");
    if (nsteps < 0) return fail(FREHD_EINVAL, "nsteps < 0");
    if (int rc = state_check(d, CALL_STEP)) return rc;
    HIPCHK(hipSetDevice(d->device));
    Ctl c;
    COPYW(d->stream, &c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
    if (c.halted || c.haltPending)
        return fail(FREHD_ECFL, "handle halted at step %d: max CFL exceeded 20", c.haltStep);
    if (c.tt + nsteps > d->cfg.Nt)
        return fail(FREHD_EINVAL, "stepping to %d exceeds the forcing series (Nt = %d)", c.tt + nsteps, d->cfg.Nt);
    d->tt = c.tt;
    Exec &e = d->self;
    for (int s = 0; s < nsteps; s++) {
        int rc = exec_step(e);
        if (rc) return rc;
        if (d->timing.on && (s % 8 == 7)) { rc = exec_collect_timing(e); if (rc) return rc; }
        if (e.hFlags[F_HALTED]) break;
    }
    int rc = exec_collect_timing(e);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(d->stream));
    COPYW(d->stream, &c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
    if (c.halted || c.haltPending) {
        double mx = 0, my = 0;
        cfl_maxima(d, &mx, &my);
        return fail(FREHD_ECFL, "max CFL exceeded 20 at step %d (CFLx %.3g, CFLy %.3g)", c.haltStep, mx, my);
    }
    return 0;
}

int frehd_dev_phase(frehd_dev_t *d, int32_t phase)
{
    if (!d) return fail(FREHD_EINVAL, "handle is NULL");
    if (int rc = state_check(d, CALL_PHASE, phase)) return rc;
    HIPCHK(hipSetDevice(d->device));
    if (d->midStep == 0 && phase >= 2) {
        // entering a step after its solve: the state's surface is the new one
        // (a state given at a boundary: no flip pending from a step run here)
        Ctl c;
        COPYW(d->stream, &c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
        if (c.flipPending) return fail(FREHD_ESTATE, "phase %d entered after a whole step: run phases 0 and 1", phase);
        const int n = d->g.Ng;
        hipLaunchKernelGGL(k_swap_surface, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->stream,
                           c.parity ? d->args.surfB : d->args.surfA, c.parity ? d->args.surfA : d->args.surfB, n);
        HIPCHK(hipGetLastError());
    }
    int rc = exec_phase(d->self, phase);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(d->stream));
    return 0;
}

int frehd_dev_download(frehd_dev_t *d, frehd_fields *o)
{
    if (!d || !o) return fail(FREHD_EINVAL, "NULL argument");
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    const size_t N = d->g.N, Ng = d->g.Ng;
    StepArgs &a = d->args;
    // the buffers holding the current depth and concentration: the device's
    // start-of-step buffers, or, between the phases of a step that is not
    // halted, the new ones once updateData (depth) or scalarTransport (S)
    // has run
    Ctl c;
    COPYW(d->stream, &c, a.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
    int dp, sp, hp;
    current_buffers(d, c, &dp, &sp, &hp);
    const double *depthCur = dp ? a.depthB : a.depthA;
    const char *Scur = reinterpret_cast<const char *>(sp ? a.S2 : a.S);
    // r: stored in the state precision
    struct { double *dst; const double *src; size_t n; bool r; } tab[] = {
        {o->surf, hp ? a.surfB : a.surfA, Ng, false}, {o->surfOld, hp ? a.surfA : a.surfB, Ng, false},
        {o->uuXP, a.uuXP, Ng, true},
        {o->vvYP, a.vvYP, Ng, true}, {o->uuYP, a.uuYP, N, true}, {o->vvXP, a.vvXP, N, true},
        {o->depth, depthCur, N, false}, {o->depthXP, a.depthXP, Ng, true}, {o->depthYP, a.depthYP, Ng, true},
        {o->EnXP, a.EnXP, Ng, true}, {o->EnYP, a.EnYP, Ng, true}, {o->DragXP, a.DragXP, Ng, true},
        {o->DragYP, a.DragYP, Ng, true}, {o->CDXP, a.CDXP, Ng, true}, {o->CDYP, a.CDYP, Ng, true},
        {o->Fuu, a.Fuu, Ng, true}, {o->Fvv, a.Fvv, Ng, true}, {o->cellV, a.cellV, N, false},
        {o->CFLx, a.CFLx, Ng, true}, {o->CFLy, a.CFLy, Ng, true}, {o->wtfXP, a.wtfXP, Ng, true},
        {o->wtfYP, a.wtfYP, Ng, true}, {o->bottomZ, a.bottomZ, Ng, false}, {o->bottomXP, a.bottomXP, Ng, false},
        {o->bottomYP, a.bottomYP, Ng, false}};
    for (auto &e : tab) {
        if (!e.dst) continue;
        if (e.r) {
            int rc = download_r(d, e.dst, e.src, e.n);
            if (rc) return rc;
        } else {
            COPYW(d->stream, e.dst, e.src, e.n * sizeof(double), hipMemcpyDeviceToHost);
        }
    }
    if (a.p.useSubgrid == 1) {
        struct { double *dst; const double *src; } sb[] = {{o->V, a.V}, {o->Z, a.Z}, {o->Vx, a.Vx},
            {o->Vy, a.Vy}, {o->Zx, a.Zx}, {o->Zy, a.Zy}, {o->Nx, a.Nx}, {o->Oy, a.Oy}};
        for (auto &e : sb) {
            int rc = download_r(d, e.dst, e.src, Ng);
            if (rc) return rc;
        }
        if (o->ind) COPYW(d->stream, o->ind, a.ind, Ng * sizeof(int), hipMemcpyDeviceToHost);
    }
    const int ns = std::max(1, a.p.nScalar);
    const size_t rw = state_f32(d->cfg) ? sizeof(float) : sizeof(double);
    for (int q = 0; q < ns; q++) {
        if (o->S) {
            int rc = download_r(d, o->S + (size_t)q * Ng,
                                reinterpret_cast<const double *>(Scur + (size_t)q * a.sStride * rw), Ng);
            if (rc) return rc;
        }
        if (o->Sm) COPYW(d->stream, o->Sm + (size_t)q * N, a.Sm + (size_t)q * a.sStride, N * sizeof(double), hipMemcpyDeviceToHost);
    }
    return 0;
}

int frehd_dev_stats(frehd_dev_t *d, frehd_stats *out)
{
    if (!d || !out) return fail(FREHD_EINVAL, "NULL argument");
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    Ctl c;
    COPYW(d->stream, &c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
    out->steps = c.steps;
    out->pcgIterTotal = c.pcgIterTotal;
    out->pcgIterLast = c.pcgIterLast;
    out->refineLast = c.refineLast;
    out->relResLast = c.relResLast;
    if (int rc = cfl_maxima(d, &out->maxCFLx, &out->maxCFLy)) return rc;
    if (int rc = vloss_totals(d, out->Vloss)) return rc;
    out->halted = c.halted || c.haltPending;
    out->haltStep = c.haltStep;
    return 0;
}

int frehd_dev_set_fused_solve(frehd_dev_t *d, int32_t max_blocks)
{
    if (!d) return fail(FREHD_EINVAL, "handle is NULL");
    if (int rc = state_check(d, CALL_TILING)) return rc;
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    d->self.fusedMax = max_blocks;
    // the launched solve's pattern history starts afresh
    d->self.mixPlan = SolveCounts();
    d->self.nRecent = 0;
    return upload_member_args(d);
}

int frehd_batch_set_fused_solve(frehd_batch_t *b, int32_t max_blocks)
{
    if (!b) return fail(FREHD_EINVAL, "batch is NULL");
    HIPCHK(hipSetDevice(b->e.device));
    HIPCHK(hipStreamSynchronize(b->e.stream));
    b->e.fusedMax = max_blocks;
    b->e.mixPlan = SolveCounts();
    b->e.nRecent = 0;
    return exec_refresh(b->e);
}

int frehd_fused_solve_capacity(int32_t cols_per_lane, int32_t *blocks)
{
    if (!blocks) return fail(FREHD_EINVAL, "blocks is NULL");
    if (cols_per_lane != 1 && cols_per_lane != 2 && cols_per_lane != 4)
        return fail(FREHD_EINVAL, "cols_per_lane %d not 1, 2 or 4", cols_per_lane);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    *blocks = fused_capacity(dev, cols_per_lane);
    return 0;
}

int frehd_dev_set_budget(frehd_dev_t *d, int32_t on)
{
    if (!d) return fail(FREHD_EINVAL, "handle is NULL");
    if (int rc = state_check(d, CALL_TILING)) return rc;
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    Budget zero;
    memset(&zero, 0, sizeof zero);
    COPYW(d->stream, &d->args.ctl->budget, &zero, sizeof zero, hipMemcpyHostToDevice);
    d->args.p.budget = on ? 1 : 0;
    return upload_member_args(d);
}

int frehd_dev_budget(frehd_dev_t *d, frehd_budget *out)
{
    if (!d || !out) return fail(FREHD_EINVAL, "NULL argument");
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    Ctl c;
    COPYW(d->stream, &c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
    memset(out, 0, sizeof *out);
    const Budget &b = c.budget;
    out->steps = b.steps;
    out->nScalar = d->args.p.nScalar;
    out->on = d->args.p.budget;
    out->volume = b.vol;
    out->volumeResidual = b.volRes;
    out->volumeResidualMax = b.volResMax;
    out->volumeBoundary = b.volBnd;
    out->volumeInflow = b.volIn;
    out->volumeClip = b.volClip;
    out->volumeMoved = b.volMoved;
    for (int q = 0; q < d->args.p.nScalar && q < FREHD_MAX_SCALARS; q++) {
        out->mass[q] = b.mass[q];
        out->massResidual[q] = b.massRes[q];
        out->massResidualMax[q] = b.massResMax[q];
        out->massBoundary[q] = b.massBnd[q];
        out->massInflow[q] = b.massIn[q];
        out->massMoved[q] = b.massMoved[q];
    }
    return 0;
}

int frehd_dev_upload(frehd_dev_t *d, const frehd_fields *in)
{
    if (!d || !in) return fail(FREHD_EINVAL, "NULL argument");
    if (int rc = state_check(d, CALL_UPLOAD)) return rc;
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    Ctl c;
    COPYW(d->stream, &c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
    // the pending buffer flips applied first, so the uploaded state is the
    // next step's start
    int dp, sp, hp;
    current_buffers(d, c, &dp, &sp, &hp);
    if (dp != c.parity || sp != c.sParity || c.flipPending || !c.runTail) {
        c.parity = dp;
        c.sParity = sp;
        c.flipPending = 0;
        c.runTail = 1;    // as a new handle: the update phase may come next
        COPYW(d->stream, d->args.ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice);
    }
    return upload_state(d, *in, dp, sp);
}

int frehd_dev_step_index(frehd_dev_t *d, int32_t *tt)
{
    if (!d || !tt) return fail(FREHD_EINVAL, "NULL argument");
    Ctl c;
    HIPCHK(hipStreamSynchronize(d->stream));
    COPYW(d->stream, &c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
    *tt = c.tt;
    return 0;
}

int frehd_dev_set_step_index(frehd_dev_t *d, int32_t tt)
{
    if (!d) return fail(FREHD_EINVAL, "handle is NULL");
    if (tt < 0 || tt > d->cfg.Nt) return fail(FREHD_EINVAL, "step index %d outside [0, Nt]", tt);
    if (int rc = state_check(d, CALL_SET_INDEX)) return rc;
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    COPYW(d->stream, &d->args.ctl->tt, &tt, sizeof(int), hipMemcpyHostToDevice);
    const int next = tt + 1;
    COPYW(d->stream, &d->args.ctl->ttNext, &next, sizeof(int), hipMemcpyHostToDevice);
    return 0;
}

int frehd_dev_destroy(frehd_dev_t *d)
{
    if (!d) return 0;
    if (d->batch) frehd_batch_remove(static_cast<frehd_batch_t *>(d->batch), d);
    hipSetDevice(d->device);
    hipStreamSynchronize(d->stream);
    exec_release_graphs(d->self);
    if (d->arena) hipFree(d->arena);
    if (d->self.hFlags) hipHostFree(d->self.hFlags);
    if (d->ownStream) hipStreamDestroy(d->stream);
    delete d;
    return 0;
}

int frehd_dev_enable_timing(frehd_dev_t *d, int32_t on)
{
    if (!d) return fail(FREHD_EINVAL, "handle is NULL");
    d->timing.on = on ? 1 : 0;
    d->timing.reset();
    return 0;
}

static int kernel_times(const Timing &t, int32_t cap, const char **names, double *ms, int64_t *launches, int32_t *n)
{
    int cnt = 0;
    for (auto &k : t.timers) {
        if (k.launches == 0) continue;
        if (cnt < cap) {
            if (names) names[cnt] = k.name;
            if (ms) ms[cnt] = k.ms / (double)k.launches;
            if (launches) launches[cnt] = k.launches;
        }
        cnt++;
    }
    *n = cnt;
    return 0;
}

int frehd_dev_kernel_times(frehd_dev_t *d, int32_t cap, const char **names, double *ms, int64_t *launches,
                           int32_t *n)
{
    if (!d || !n) return fail(FREHD_EINVAL, "NULL argument");
    return kernel_times(d->timing, cap, names, ms, launches, n);
}

int frehd_dev_download_tables(frehd_dev_t *d, double *const tables[10])
{
    if (!d || !tables) return fail(FREHD_EINVAL, "NULL argument");
    if (d->cfg.useSubgrid != 1) return fail(FREHD_ESTATE, "the handle has no subgrid tables");
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    const size_t tn = (size_t)d->cfg.Ne * d->g.N;
    const StepArgs &a = d->args;
    // on the fly: generate the interleaved tables from the pixels first, in
    // the precision the handle's lookups round them to
    const int f32 = state_f32(d->cfg);
    const size_t tw = f32 ? sizeof(float) : sizeof(double);
    void *gen = nullptr;
    if (!a.tabs) {
        if (hipMalloc(&gen, tn * 10 * tw) != hipSuccess)
            return fail(FREHD_ENOMEM, "hipMalloc of the generated tables (%zu bytes) failed", tn * 10 * tw);
        if (launch_subgrid_tables(a.bed.z, gen, d->cfg.nx, d->cfg.ny, a.bed.r, d->cfg.Ne, d->cfg.surfmin,
                                  d->cfg.dsurf, a.bed.dxf, f32, d->stream)) {
            hipFree(gen);
            return fail(FREHD_EHIP, "table generator launch failed");
        }
    }
    const void *src = a.tabs ? static_cast<const void *>(a.tabs) : gen;
    double *stage = nullptr;
    if (hipMalloc(&stage, tn * sizeof(double)) != hipSuccess) {
        hipStreamSynchronize(d->stream);
        if (gen) hipFree(gen);
        return fail(FREHD_ENOMEM, "hipMalloc of the table staging buffer failed");
    }
    int rc = 0;
    for (int t = 0; t < 10 && !rc; t++) {
        if (!tables[t]) continue;
        if (launch_tables_reorder(const_cast<void *>(src), stage, (int)d->g.N, d->cfg.Ne, t, 0, f32, d->stream))
            rc = fail(FREHD_EHIP, "table reorder launch failed");
        else if (hipMemcpyAsync(tables[t], stage, tn * sizeof(double), hipMemcpyDeviceToHost, d->stream) != hipSuccess ||
                 hipStreamSynchronize(d->stream) != hipSuccess)
            rc = fail(FREHD_EHIP, "table download failed");
    }
    hipStreamSynchronize(d->stream);
    hipFree(stage);
    if (gen) hipFree(gen);
    return rc;
}

int frehd_dev_set_pcg_tiling(frehd_dev_t *d, int32_t rows_per_tile, int32_t xcd_remap, int32_t cols_per_lane)
{
    if (!d) return fail(FREHD_EINVAL, "handle is NULL");
    if (rows_per_tile < 0 || rows_per_tile > d->g.ny)
        return fail(FREHD_EINVAL, "rows_per_tile %d out of range [0, %d]", rows_per_tile, d->g.ny);
    if (cols_per_lane != 1 && cols_per_lane != 2 && cols_per_lane != 4)
        return fail(FREHD_EINVAL, "cols_per_lane must be 1, 2 or 4");
    if (cols_per_lane > max_lanes(d->cfg) || d->g.nx % cols_per_lane)
        return fail(FREHD_EINVAL, "cols_per_lane %d: at most %d for this solver (4 mixed, 2 FP64 Jacobi, 1 SSOR), "
                    "and nx a multiple of it", cols_per_lane, max_lanes(d->cfg));
    if ((size_t)dir_tiling(d->g.nx, d->g.ny, rows_per_tile, xcd_remap).nb > max_dir_blocks(d->g.nx, d->g.ny))
        return fail(FREHD_EINVAL, "too many tiles");
    if (int rc = state_check(d, CALL_TILING)) return rc;
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    set_pcg_args(d, rows_per_tile, xcd_remap ? 1 : 0, cols_per_lane);
    return upload_member_args(d);
}

int frehd_dev_autotune(frehd_dev_t *d, int32_t iters, int32_t *rows_per_tile, int32_t *xcd_remap,
                       int32_t *cols_per_lane, double *ms_per_iteration)
{
    if (!d) return fail(FREHD_EINVAL, "handle is NULL");
    if (int rc = state_check(d, CALL_AUTOTUNE)) return rc;
    return autotune_group(d->self, iters, rows_per_tile, xcd_remap, cols_per_lane, ms_per_iteration);
}

// ---- batches ----------------------------------------------------------------
int frehd_batch_create(int32_t capacity, frehd_batch_t **out)
{
    if (!out) return fail(FREHD_EINVAL, "out is NULL");
    if (capacity < 1 || capacity > 65535) return fail(FREHD_EINVAL, "capacity %d outside [1, 65535]", capacity);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FREHD_EHIP, "no HIP device visible");
    frehd_batch *b = new frehd_batch();
    Exec &e = b->e;
    e.capacity = capacity;
    HIPCHK(hipGetDevice(&e.device));
    HIPCHK(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
    const size_t n = (size_t)capacity;
    if (hipMalloc(&e.dArgs, n * sizeof(StepArgs)) != hipSuccess ||
        hipMalloc(&e.dPcg, n * sizeof(PcgArgs<double>)) != hipSuccess ||
        hipMalloc(&e.dPcg32, n * sizeof(PcgArgs<float>)) != hipSuccess ||
        hipMalloc(&e.dRef, n * sizeof(RefineArgs)) != hipSuccess ||
        hipMalloc(&e.dSync, sizeof(SolveSync)) != hipSuccess ||
        hipMalloc(&e.dPlan, kPlanCap * sizeof(SolveBlock)) != hipSuccess ||
        alloc_flags(e, n) != hipSuccess) {
        frehd_batch_destroy(b);
        return fail(FREHD_ENOMEM, "batch allocation failed");
    }
    e.planCap = kPlanCap;
    HIPCHK(hipMemsetAsync(e.dSync, 0, sizeof(SolveSync), e.stream));
    HIPCHK(hipStreamSynchronize(e.stream));
    e.tim = &b->timing;
    e.useGraphs = graphs_from_env();
    e.patternCap = pattern_cap_from_env();
    *out = b;
    return 0;
}

int frehd_batch_add(frehd_batch_t *b, frehd_dev_t *d)
{
    if (!b || !d) return fail(FREHD_EINVAL, "NULL argument");
    if (d->batch) return fail(FREHD_ESTATE, "the handle is already in a batch");
    if (d->midStep) return fail(FREHD_ESTATE, "a step of the handle is half done (%d of its 5 phases ran)", d->midStep);
    if (d->device != b->e.device) return fail(FREHD_EINVAL, "the handle lives on another device");
    if ((int)b->e.members.size() >= b->e.capacity) return fail(FREHD_EBUDGET, "batch full (%d)", b->e.capacity);
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    // the solver's launches are shared: the newcomer takes the members'
    // columns per lane, or everyone goes back to one if its nx does not
    // allow that
    auto lanes_of = [](const frehd_dev_impl *m) {
        return m->cfg.precision == FREHD_PREC_MIXED ? m->pcg32.lanes : m->pcg.lanes;
    };
    if (!b->e.members.empty() && max_lanes(d->cfg) > 1 && max_lanes(b->e.members[0]->cfg) > 1) {
        const int L = lanes_of(b->e.members[0]);
        if (lanes_of(d) != L) {
            int rc = 0;
            if (d->g.nx % L == 0) {
                set_pcg_args(d, d->tileH, d->tileRemap, L);
                rc = upload_member_args(d);
            } else {
                for (auto *m : b->e.members) {
                    frehd_dev *md = static_cast<frehd_dev *>(m);
                    set_pcg_args(md, md->tileH, md->tileRemap, 1);
                    if (!rc) rc = upload_member_args(md);
                }
                if (!rc && lanes_of(d) != 1) {
                    set_pcg_args(d, d->tileH, d->tileRemap, 1);
                    rc = upload_member_args(d);
                }
            }
            if (rc) return rc;
        }
    }
    // ... and the members' block shape of the stencil kernels
    if (!b->e.members.empty() && d->blockX != b->e.members[0]->blockX) {
        d->blockX = b->e.members[0]->blockX;
        fill_extents(d);
        if (int rc = upload_member_args(d)) return rc;
    }
    b->e.members.push_back(d);
    int rc = exec_refresh(b->e);
    if (rc) {
        b->e.members.pop_back();
        exec_refresh(b->e);
        return rc;
    }
    d->batch = b;
    return 0;
}

int frehd_batch_remove(frehd_batch_t *b, frehd_dev_t *d)
{
    if (!b || !d) return fail(FREHD_EINVAL, "NULL argument");
    auto &m = b->e.members;
    auto it = std::find(m.begin(), m.end(), static_cast<frehd_dev_impl *>(d));
    if (it == m.end()) return fail(FREHD_EINVAL, "the handle is not in this batch");
    HIPCHK(hipStreamSynchronize(b->e.stream));
    m.erase(it);
    d->batch = nullptr;
    return exec_refresh(b->e);
}

int frehd_batch_size(frehd_batch_t *b, int32_t *n)
{
    if (!b || !n) return fail(FREHD_EINVAL, "NULL argument");
    *n = (int32_t)b->e.members.size();
    return 0;
}

int frehd_batch_step(frehd_batch_t *b, int32_t nsteps)
{
    if (!b) return fail(FREHD_EINVAL, "batch is NULL");
    if (nsteps < 0) return fail(FREHD_EINVAL, "nsteps < 0");
    Exec &e = b->e;
    if (e.members.empty() || nsteps == 0) return 0;
    HIPCHK(hipSetDevice(e.device));
    for (auto *d : e.members) {
        Ctl c;
        COPYW(d->stream, &c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
        if (!(c.halted || c.haltPending) && c.tt + nsteps > d->cfg.Nt)
            return fail(FREHD_EINVAL, "a member steps to %d beyond its forcing series (Nt = %d)", c.tt + nsteps,
                        d->cfg.Nt);
        d->tt = c.tt;
    }
    for (int s = 0; s < nsteps; s++) {
        int rc = exec_step(e);
        if (rc) return rc;
        if (b->timing.on && (s % 8 == 7)) { rc = exec_collect_timing(e); if (rc) return rc; }
    }
    int rc = exec_collect_timing(e);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e.stream));
    int halted = 0;
    for (auto *d : e.members) {
        Ctl c;
        COPYW(d->stream, &c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
        halted += (c.halted || c.haltPending) ? 1 : 0;
    }
    if (halted) return fail(FREHD_ECFL, "%d batch member(s) halted: max CFL exceeded 20", halted);
    return 0;
}

int frehd_batch_autotune(frehd_batch_t *b, int32_t iters, int32_t *rows_per_tile, int32_t *xcd_remap,
                         int32_t *cols_per_lane, double *ms_per_iteration)
{
    if (!b) return fail(FREHD_EINVAL, "batch is NULL");
    int rc = autotune_group(b->e, iters, rows_per_tile, xcd_remap, cols_per_lane, ms_per_iteration);
    // the members' own groups carry the new tiling too
    for (auto *d : b->e.members)
        if (!rc) rc = upload_member_args(d);
    return rc;
}

// Block shape of the 2-D stencil kernels chosen by timing whole steps: each
// candidate (64 x 4, 128 x 2, 256 x 1 cells) runs `steps` steps twice, in two
// passes, with per-kernel event timing; the one whose predictor, limiter
// and scalar kernels take the least time per step wins.  step(k) advances
// the group k steps.  The shape changes no cell's arithmetic (only the order
// of the block partial sums of the volume losses and budgets), so the model
// advances as it would with any shape.
static int autotune_blocks(Exec &e, std::vector<frehd_dev_impl *> members, int steps, int32_t *bestX, double *bestMs,
                           const std::function<int(int)> &step)
{
    if (members.empty()) return fail(FREHD_ESTATE, "nothing to tune");
    if (steps < 1 || steps > 1000) return fail(FREHD_EINVAL, "steps %d outside [1, 1000]", steps);
    for (auto *d : members) {
        Ctl c;
        COPYW(d->stream, &c, d->args.ctl, sizeof(Ctl), hipMemcpyDeviceToHost);
        if (!(c.halted || c.haltPending) && c.tt + 6 * steps > d->cfg.Nt)
            return fail(FREHD_EINVAL, "tuning steps to %d beyond the forcing series (Nt = %d)", c.tt + 6 * steps,
                        d->cfg.Nt);
    }
    Timing local;
    local.on = 1;
    struct Swap {
        Exec &e;
        Timing *saved;
        ~Swap() { e.tim = saved; }
    } swap{e, e.tim};
    e.tim = &local;
    const int cands[3] = {64, 128, 256};
    double ms[3] = {-1.0, -1.0, -1.0};
    int rc = 0;
    for (int pass = 0; pass < 2 && !rc; pass++)
        for (int q = 0; q < 3 && !rc; q++) {
            for (auto *d : members) {
                frehd_dev *dd = static_cast<frehd_dev *>(d);
                dd->blockX = cands[q];
                fill_extents(dd);
                if ((rc = upload_member_args(dd))) break;
            }
            if (!rc && &e != &members[0]->self) rc = exec_refresh(e);
            if (rc) break;
            local.reset();
            if ((rc = step(steps))) break;
            double t = 0.0;
            for (int k : {K_PRED, K_LIM, K_SCALAR})
                if ((size_t)k < local.timers.size()) t += local.timers[k].ms;
            t /= steps;
            if (t > 0 && (ms[q] < 0 || t < ms[q])) ms[q] = t;
        }
    int best = 0;
    for (int q = 1; q < 3; q++)
        if (ms[q] > 0 && ms[q] < ms[best]) best = q;
    for (auto *d : members) {
        frehd_dev *dd = static_cast<frehd_dev *>(d);
        dd->blockX = rc ? 64 : cands[best];
        fill_extents(dd);
        int r2 = upload_member_args(dd);
        if (!rc) rc = r2;
    }
    if (!rc && &e != &members[0]->self) rc = exec_refresh(e);
    if (rc) return rc;
    if (bestX) *bestX = cands[best];
    if (bestMs) *bestMs = ms[best];
    return 0;
}

int frehd_dev_autotune_blocks(frehd_dev_t *d, int32_t steps, int32_t *block_x, double *ms_per_step)
{
    if (!d) return fail(FREHD_EINVAL, "handle is NULL");
    if (int rc = state_check(d, CALL_AUTOTUNE)) return rc;
    HIPCHK(hipSetDevice(d->device));
    HIPCHK(hipStreamSynchronize(d->stream));
    return autotune_blocks(d->self, {d}, steps, block_x, ms_per_step, [&](int k) {
        int rc = frehd_dev_step(d, k);
        return rc == FREHD_ECFL ? 0 : rc;
    });
}

int frehd_batch_autotune_blocks(frehd_batch_t *b, int32_t steps, int32_t *block_x, double *ms_per_step)
{
    if (!b) return fail(FREHD_EINVAL, "batch is NULL");
    HIPCHK(hipSetDevice(b->e.device));
    HIPCHK(hipStreamSynchronize(b->e.stream));
    return autotune_blocks(b->e, b->e.members, steps, block_x, ms_per_step, [&](int k) {
        int rc = frehd_batch_step(b, k);
        return rc == FREHD_ECFL ? 0 : rc;
    });
}

int frehd_batch_enable_timing(frehd_batch_t *b, int32_t on)
{
    if (!b) return fail(FREHD_EINVAL, "batch is NULL");
    b->timing.on = on ? 1 : 0;
    b->timing.reset();
    return 0;
}

int frehd_batch_kernel_times(frehd_batch_t *b, int32_t cap, const char **names, double *ms, int64_t *launches,
                             int32_t *n)
{
    if (!b || !n) return fail(FREHD_EINVAL, "NULL argument");
    return kernel_times(b->timing, cap, names, ms, launches, n);
}

int frehd_batch_destroy(frehd_batch_t *b)
{
    if (!b) return 0;
    Exec &e = b->e;
    hipSetDevice(e.device);
    if (e.stream) hipStreamSynchronize(e.stream);
    for (auto *d : e.members) d->batch = nullptr;
    exec_release_graphs(e);
    if (e.dArgs) hipFree(e.dArgs);
    if (e.dPcg) hipFree(e.dPcg);
    if (e.dPcg32) hipFree(e.dPcg32);
    if (e.dRef) hipFree(e.dRef);
    if (e.dSync) hipFree(e.dSync);
    if (e.dPlan) hipFree(e.dPlan);
    if (e.hFlags) hipHostFree(e.hFlags);
    if (e.stream) hipStreamDestroy(e.stream);
    delete b;
    return 0;
}

}  // extern "C"
