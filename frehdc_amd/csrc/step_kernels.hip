// step_kernels.hip -- FP64 kernels of one Casulli time step (plain grid).
//
// Each kernel covers a fused group of the reference's per-cell sweeps; every
// arithmetic expression keeps the reference's evaluation order and the file
// is compiled with -ffp-contract=off, so on identical inputs each kernel
// reproduces the reference bit for bit (pow() of the drag update aside).
// Sequential in-place sweeps of the reference are restated as per-cell
// closed forms (see the comment at each kernel).
#include "frehd_dev.h"
#include "step_kernels.h"

namespace frehd {

template <class R>
__device__ inline int cur_tt(const StepArgsT<R> &a) { return a.ctl->tt; }

// ---------------------------------------------------------------------------
// Step begin: tt += 1 (SolveAll's loop counter, nssolve.c:61) and the solver
// state reset; done by the predictor's first thread (the rest of the
// predictor reads the step index from ttNext, not tt), so a step costs no
// launch for it.
template <class R>
__device__ inline void step_begin(const StepArgsT<R> &a)
{
    Ctl *c = a.ctl;
    if (c->halted) return;
    if (c->haltPending) {
        // monitorCFL's stop after the last step (nsfunctions.c:1226-1231):
        // nothing of this step runs; that step's buffer flips stay pending
        // (the host reads its state through them)
        c->halted = 1;
        return;
    }
    if (c->flipPending) {
        c->parity ^= 1;                       // the new depth becomes this step's start
        if (a.p.nScalar > 0) c->sParity ^= 1; // and the new concentration
        c->flipPending = 0;
    }
    c->tt += 1;
    c->runTail = 0;
    c->tailIssued = 0;
    c->pcg.iter = 0;
    c->pcg.doneAt = -1;
    c->ref.doneAt = -1;
    c->ref.sweep = 0;
    c->ref.innerConv = 0;
    c->pcgIn.doneAt = -1;
    c->pcgIn.iter = 0;
    c->pcg.rhoHist[0] = c->pcg.rhoHist[1] = 0.0;
}

// monitorCFL's stop test (nsfunctions.c:1226-1231) on one block's maxima of
// the new CFL numbers: a block whose cells exceed 20 marks the member (the
// global maximum exceeds 20 exactly when some block's does); the next
// step's begin halts it, since the reference completes the step and then
// exits.  Any number of blocks may write the same values.
template <class R>
__device__ inline void cfl_stop_test(const StepArgsT<R> &a, double mx, double my)
{
    if (fabs(mx) > 20 || fabs(my) > 20) {
        a.ctl->haltPending = 1;
        a.ctl->haltStep = a.ctl->tt;
    }
}

// ---------------------------------------------------------------------------
// Predictor: explicitVelocity + advectionTerm + diffusionTerm + dragTerm +
// dragInversion (nsfunctions.c:498-746).
//
// advectionTerm's CFL limiter walks cells in increasing index and rescales
// the advection term of the cell and of its four neighbours in place, and
// each cell's En is updated right after its own turn (nsfunctions.c:536-585).
// So the term of cell c that enters En[c] has been rescaled by the cells
// processed up to c that list c as a target: c - nx (c is its north
// neighbour), c - 1 (when c is its east neighbour, i.e. i > 0) and c itself,
// in that order.  Each rescale multiplies by (CFLh - CFL[c])/(CFLh - CFLl)
// or zeroes, depending on the CFL of the cell doing the rescaling.
__device__ inline void limit_x(double &aX, double cflK, double fX, double lo, double hi, double den)
{
    if (cflK > lo && cflK <= hi) aX = aX * fX / den;
    else if (cflK > hi) aX = 0;
}
__device__ inline void limit_y(double &aY, double cflYK, double cflXK, double fY, double lo, double hi,
                               double den)
{
    if (cflYK > lo && cflYK < hi) aY = aY * fY / den;
    else if (cflXK > hi) aY = 0;   // tests CFLx, nsfunctions.c:574
}

// the predictor's results of one cell: En and Drag of its east and north faces
struct Pred {
    double EnX, EnY, DragX, DragY;
};

template <class R, bool SUB>
__device__ inline Pred predict(const StepArgsT<R> &a, int c, int i, int j)
{
    const Grid &g = a.g;
    const int e = nbE(g, c, i, j), w = nbW(g, c, i, j), n = nbN(g, c, i, j), s = nbS(g, c, i, j);
    const P &p = a.p;
    const double dx = p.dx, dy = p.dy, dt = p.dt;
    const auto u = a.uuXP;
    const auto v = a.vvYP;
    const double uc = u[c], ue = u[e], uw = u[w], un = u[n], us = u[s];
    const double vc = v[c], ve = v[e], vw = v[w], vn = v[n], vs = v[s];
    const double vx = a.vvXP[c], uy = a.uuYP[c];

    double aX = (0.5 / dx) * (uc + fabs(uc)) * (uc - uw) + (0.5 / dx) * (uc - fabs(uc)) * (ue - uc) +
                (0.5 / dy) * (vx + fabs(vx)) * (uc - us) + (0.5 / dy) * (vx - fabs(vx)) * (un - uc);
    double aY = (0.5 / dx) * (uy + fabs(uy)) * (vc - vw) + (0.5 / dx) * (uy - fabs(uy)) * (ve - vc) +
                (0.5 / dy) * (vc + fabs(vc)) * (vc - vs) + (0.5 / dy) * (vc - fabs(vc)) * (vn - vc);
    if (uc == 0) aX = 0;
    if (vc == 0) aY = 0;

    const double lo = p.CFLl, hi = p.CFLh, den = hi - lo;
    const double cxc = a.CFLx[c], cyc = a.CFLy[c];
    const double fX = hi - cxc, fY = hi - cyc;
    if (j > 0) {
        const double kx = a.CFLx[c - g.nx], ky = a.CFLy[c - g.nx];
        limit_x(aX, kx, fX, lo, hi, den);
        limit_y(aY, ky, kx, fY, lo, hi, den);
    }
    if (i > 0) {
        const double kx = a.CFLx[c - 1], ky = a.CFLy[c - 1];
        limit_x(aX, kx, fX, lo, hi, den);
        limit_y(aY, ky, kx, fY, lo, hi, den);
    }
    limit_x(aX, cxc, fX, lo, hi, den);
    limit_y(aY, cyc, cxc, fY, lo, hi, den);

    double EnX = uc - dt * aX;
    double EnY = vc - dt * aY;

    double dfx, dfy;
    if (SUB) {
        // diffusionTermSub, subgrid.c:781-815
        const double Vxc = a.Vx[c], Vyc = a.Vy[c];
        const double Nxc = a.Nx[c], Nxe = a.Nx[e], Nxw = a.Nx[w];
        const double Oyc = a.Oy[c], Oyn = a.Oy[n], Oys = a.Oy[s];
        dfx = 0;
        dfy = 0;
        if (Vxc > 0)
            dfx = (p.NUx / Vxc) * (Nxe * (ue - uc) / dx - Nxc * (-uw + uc) / dx) +
                  (p.NUy / Vxc) * (Oyc * (un - uc) / dy - Oys * (-us + uc) / dy);
        if (Vyc > 0)
            dfy = (p.NUx / Vyc) * (Nxc * (ve - vc) / dx - Nxw * (-vw + vc) / dx) +
                  (p.NUy / Vyc) * (Oyn * (vn - vc) / dy - Oyc * (-vs + vc) / dy);
    } else {
        dfx = (p.NUx / (dx * dx)) * (ue - 2 * uc + uw) + (p.NUy / (dy * dy)) * (us - 2 * uc + un);
        dfy = (p.NUx / (dx * dx)) * (ve - 2 * vc + vw) + (p.NUy / (dy * dy)) * (vs - 2 * vc + vn);
    }
    EnX = EnX + dt * dfx;
    EnY = EnY + dt * dfy;

    // windTerm (nsfunctions.c:609-648; the plain term on the subgrid path
    // too, nssolve.c:131): stress rhoa Cw |W - u.e|^2 along the wind
    // direction, decaying on thin faces, over the face depth
    if (p.useWind == 1) {
        const int tt = a.ctl->ttNext;
        const double cw = a.windCos[tt], sw = a.windSin[tt], spd = a.windSpd[tt];
        const double rel = spd - uc * cw - vc * sw;
        const double tau = p.rhoa * p.Cw * rel * rel;
        const double dX = a.depthXP[c], dY = a.depthYP[c];
        const double tX = (dX < p.hD && p.useThinLayer == 1) ? tau * exp(p.CwT * (dX - p.hD) / p.hD) : tau;
        const double tY = (dY < p.hD && p.useThinLayer == 1) ? tau * exp(p.CwT * (dY - p.hD) / p.hD) : tau;
        if (dX > 0) EnX = EnX + dt * tX * cw / (dX * 1000.0);
        if (dY > 0) EnY = EnY - dt * tY * sw / (dY * 1000.0);
    }

    const double velx = sqrt(uc * uc + vx * vx);
    const double vely = sqrt(uy * uy + vc * vc);
    double DragX, DragY;
    if (SUB) {
        // dragTermSub with staggeredV = 1, subgrid.c:857-927
        const double Vxc = a.Vx[c], Vyc = a.Vy[c];
        const double effHx = Vxc > 0 ? a.Zx[c] / Vxc : 0;
        const double effHy = Vyc > 0 ? a.Zy[c] / Vyc : 0;
        DragX = 0.5 * dt * a.CDXP[c] * velx * effHx;
        DragY = 0.5 * dt * a.CDYP[c] * vely * effHy;
    } else {
        const double dX = a.depthXP[c], dY = a.depthYP[c];
        DragX = dX > 0 ? 0.5 * dt * a.CDXP[c] * velx / dX : 0;
        DragY = dY > 0 ? 0.5 * dt * a.CDYP[c] * vely / dY : 0;
    }
    EnX = EnX / (1 + DragX);
    EnY = EnY / (1 + DragY);
    return Pred{EnX, EnY, DragX, DragY};
}

// the values as the state stores them (rounded to R): what a later kernel
// reads back
template <class R>
__device__ inline Pred stored(const Pred &q)
{
    return Pred{(double)(R)q.EnX, (double)(R)q.EnY, (double)(R)q.DragX, (double)(R)q.DragY};
}

template <class R>
__device__ inline void store_pred(const StepArgsT<R> &a, int c, int i, int j, const Pred &q)
{
    const Grid &g = a.g;
    a.EnXP[c] = q.EnX;
    a.EnYP[c] = q.EnY;
    a.DragXP[c] = q.DragX;
    a.DragYP[c] = q.DragY;
    // ghost copies (dragTerm :679-686, dragInversion :739-745)
    if (j == 0) { a.DragXP[gjM(g, i)] = q.DragX; a.EnXP[gjM(g, i)] = q.EnX; }
    if (i == 0) { a.DragYP[giM(g, j)] = q.DragY; a.EnYP[giM(g, j)] = q.EnY; }
}

template <class R, bool SUB>
__device__ inline void predictor_body(const StepArgsT<R> &a)
{
    if (a.ctl->halted || a.ctl->haltPending) return;
    const Grid g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = row_block(a) * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny) return;
    const int c = j * g.nx + i;
    store_pred(a, c, i, j, predict<R, SUB>(a, c, i, j));
}

template <class R>
__global__ void __launch_bounds__(256) k_predictor(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.gx2 || blockIdx.y >= a.gy2) return;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) step_begin(a);
    predictor_body<R, false>(a);
}
// (at least 6 waves per SIMD: the compiler then keeps it to 72 registers, 7
// waves, with no spill; by default it takes 86, 5 waves)
template <class R>
__global__ void __launch_bounds__(256, 6) k_predictor_sub(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.gx2 || blockIdx.y >= a.gy2) return;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) step_begin(a);
    predictor_body<R, true>(a);
}

// ---------------------------------------------------------------------------
// Assembly: matrixSourceTerm + matrixCoeff + adjustMatrixBoundary +
// adjustTidalBoundary (nsfunctions.c:767-924) into the matrix-free operator,
// plus the PCG start (solveMatrix :1004-1017: x = 0, r = b).
//
// Operator storage per cell: cX = G of the east face, cY = G of the north
// face, D = GnCt, with D = -1 marking a Dirichlet (tidal) row, whose matrix
// row is the identity.  Faces on the domain walls have zero depth, so the
// wall adjustments of adjustMatrixBoundary subtract zeros and GnCt comes out
// as (((1 + GnXP) + GnYP) + GnXM) + GnYM with GnXM = cX[west], GnYM = cY[south].
__device__ inline double face_coef(double coef, double depthF, double dragF)
{
    return coef * depthF / (1 + dragF);
}

// one operator row: G of the east and north faces, the centre (D = -1: a
// Dirichlet row) and the right-hand side
struct Row {
    double GXP, GYP, D, z;
};

// stores the operator row and the solver start of cell c, accumulating r.r,
// r.z (z = r / |D|) and b.b of the thread's cells (r = b from a zero start)
template <class R>
__device__ inline void store_row(const StepArgsT<R> &a, int c, const Row &w, double (&part)[3])
{
    const double D = w.D, z = w.z;
    a.cX[c] = w.GXP;
    a.cY[c] = w.GYP;
    a.D[c] = D;
    a.b[c] = z;
    // the solver's start: 0 (the reference's), the current surface, or its
    // linear extrapolation in time 2 eta^n - eta^(n-1).  From a nonzero
    // start the solver forms r = b - A x itself (the mixed path's refinement,
    // k_pcg1_start for FP64).
    a.x[c] = a.p.warmStart == 2 ? 2.0 * sf_start(a)[c] - sf_new(a)[c] : (a.p.warmStart ? sf_start(a)[c] : 0.0);
    if (a.cX32 || a.p.warmStart) return;
    a.r[c] = z;
    const double invd = 1.0 / fabs(D);
    part[0] += z * z;
    part[1] += z * (z * invd);
    part[2] += z * z;
}

// the predictor's values one row needs: En, Drag of the cell's east / north
// faces (c) and of its west (w) and south (s) neighbours' faces
struct RowEn {
    double EnXc, EnXw, EnYc, EnYs, DXc, DXw, DYc, DYs;
};

// one row of the plain operator and its right-hand side (the body of the
// reference's loops, in their order); tt: the step being assembled
template <class R>
__device__ inline Row assemble_row(const StepArgsT<R> &a, int c, int i, int j, int tt, const RowEn &q)
{
    const Grid &g = a.g;
    const P &p = a.p;
    const double dt = p.dt;
    const double coefx = p.g * (dt * dt / (p.dx * p.dx));
    const double coefy = p.g * (dt * dt / (p.dy * p.dy));
    const int w = nbW(g, c, i, j), s = nbS(g, c, i, j);
    const double GXP = face_coef(coefx, a.depthXP[c], q.DXc);
    const double GXM = face_coef(coefx, a.depthXP[w], q.DXw);
    const double GYP = face_coef(coefy, a.depthYP[c], q.DYc);
    const double GYM = face_coef(coefy, a.depthYP[s], q.DYs);
    double D = 1 + GXP + GYP + GXM + GYM;

    double z = sf_start(a)[c] - (dt / p.dx) * (q.EnXc * a.depthXP[c] - q.EnXw * a.depthXP[w]) -
               (dt / p.dy) * (q.EnYc * a.depthYP[c] - q.EnYs * a.depthYP[s]);
    const uint8_t fl = a.flags[c];
    const int ninflow = fl >> 2;
    for (int k = 0; k < ninflow; k++)
        z = z + (dt / (p.dx * p.dy)) * (a.inflow[tt] / p.inflowLocLength);
    if ((fl & 1) && p.bcType != 2) { z = a.tideP[tt] + p.offset; D = -1.0; }
    if ((fl & 2) && p.bcType != 1) { z = a.tideM[tt] + p.offset; D = -1.0; }
    return Row{GXP, GYP, D, z};
}

// ---------------------------------------------------------------------------
// Surface update: getFreeSurface + zeroNegSurf + enforceFreeSurfaceBC +
// enforceTidalBC (nsfunctions.c:1020-1031, 436-449, 317-334, 412-433), over
// all Ng indices.  A ghost takes the zero-clipped solution of the interior
// cell it mirrors (the tidal override comes after the ghost fill).
template <class R>
__global__ void __launch_bounds__(256) k_surface(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.nb1) return;
    __shared__ double lds[16];
    if (tail_off(a)) return;
    const Grid g = a.g;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double loss[1] = {0.0};
    // the solution: x, or for the mixed path the refinement buffer of the
    // sweep that converged
    const gptr<const double> xs = (a.cX32 && (a.ctl->ref.doneAt & 1)) ? a.xB : a.x;
    if (k < g.Ng) surface_cell(a, k, loss[0], xs);
    block_sum<1>(loss, lds);
    if (threadIdx.x == 0) a.vpart[2 * a.maxBlocks + blockIdx.x] += loss[0];
}


// ---------------------------------------------------------------------------
// limiterCFL + computeVolume + updateDepth (nsfunctions.c:452-495, 361-370,
// 239-247).  limiterCFL's first loop writes only surf[c] of the cell tested
// and reads surfOld and depth (old) of its neighbours, so it is per-cell; the
// new depth goes to a second buffer because neighbours read the old one.
// depth at a ghost index is 0 (the reference reads past the N-array).
// limiterCFL for cell c (both sweeps); returns the new surface, sets rmv and
// adds the Vloss1/Vloss2 contributions
template <class R>
__device__ inline double limiter_cell(const StepArgsT<R> &a, int c, int i, int j, double h, double (&loss)[2],
                                      uint8_t &rmv)
{
    const Grid &g = a.g;
    const P &p = a.p;
    const double *dOld = depth_cur(a), *ho = sf_start(a);
    const double zb = a.bottomZ[c];
    double diff = h - zb;
    rmv = 0;
    if (dOld[c] <= 0 && diff > 0) {
        const double sref = ho[c] + p.minDepth;
        const int e = nbE(g, c, i, j), w = nbW(g, c, i, j), n = nbN(g, c, i, j), s = nbS(g, c, i, j);
        const bool dE = ho[e] <= sref || (e < g.N ? dOld[e] : 0.0) <= 0;
        const bool dW = ho[w] <= sref || (w < g.N ? dOld[w] : 0.0) <= 0;
        const bool dN = ho[n] <= sref || (n < g.N ? dOld[n] : 0.0) <= 0;
        const bool dS = ho[s] <= sref || (s < g.N ? dOld[s] : 0.0) <= 0;
        if (dE && dW && dN && dS) {
            if ((double)c < (double)g.N - p.dx) loss[0] += diff * p.dx * p.dy;
            h = zb;
            rmv = 1;
        }
    }
    diff = h - zb;
    if (diff <= p.minDepth && diff > 0) {
        loss[1] += diff * p.dx * p.dy;
        h = zb;
        rmv = 0;
    }
    return h;
}

// the volume losses of a 2-D block into its accumulator slots: Vloss1,
// Vloss2 (limiterCFL) and, where the kernel also made the surface update,
// Vloss3 (zeroNegSurf)
template <class R, int NL>
__device__ inline void store_vloss(const StepArgsT<R> &a, double (&loss)[NL], double *lds)
{
    block_sum<NL>(loss, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const int bid = blockIdx.x + a.gx2 * row_block(a);
#pragma unroll
        for (int q = 0; q < NL; q++) a.vpart[(size_t)q * a.maxBlocks + bid] += loss[q];
    }
}

// The surface update (surface_cell) of interior cell c made inside the
// limiter kernel: c's new surface (tidal values, clip loss into loss), and
// the ghost entries that mirror c -- its row's jP / jM ghost, its column's
// iP / iM ghost and the corners (ghost_source) -- written here.  The
// limiter reads the neighbours' old surface from sf_start, so no thread
// depends on another's new value.
template <class R>
__device__ inline double fused_surface(const StepArgsT<R> &a, int c, int i, int j, double &loss)
{
    const Grid &g = a.g;
    const gptr<const double> xs = (a.cX32 && (a.ctl->ref.doneAt & 1)) ? a.xB : a.x;
    const gptr<double> hn = sf_new(a);
    const bool top = j == g.ny - 1, bot = j == 0, right = i == g.nx - 1, left = i == 0;
    if (top || bot || right || left) {
        double none = 0.0;
        const double hg = surface_value(a, gjP(g, i), c, none, xs);   // any ghost of c: the clipped x of c
        if (top) hn[gjP(g, i)] = hg;
        if (bot) hn[gjM(g, i)] = hg;
        if (right) hn[giP(g, j)] = hg;
        if (left) hn[giM(g, j)] = hg;
        if (top && right) hn[cIPJP(g)] = hg;
        if (top && left) hn[cIPJM(g)] = hg;
        if (bot && right) hn[cIMJP(g)] = hg;
        if (bot && left) hn[cIMJM(g)] = hg;
    }
    return surface_value(a, c, -1, loss, xs);
}

// FUSED: the surface update is made here too (the step's separate surface
// sweep, k_surface, is left out)
template <class R, bool FUSED>
__global__ void __launch_bounds__(256) k_limiter(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.gx2 || blockIdx.y >= a.gy2) return;
    __shared__ double lds[16 * 3];
    if (tail_off(a)) return;
    // the new depth goes to the second buffer: the next step starts from it
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) a.ctl->flipPending = 1;
    const Grid g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = row_block(a) * blockDim.y + threadIdx.y;
    double loss[3] = {0.0, 0.0, 0.0};
    double (&lim)[2] = *reinterpret_cast<double (*)[2]>(loss);
    if (i < g.nx && j < g.ny) {
        const P &p = a.p;
        const int c = j * g.nx + i;
        uint8_t rmv;
        const double hs = FUSED ? fused_surface(a, c, i, j, loss[2]) : sf_new(a)[c];
        const double h = limiter_cell(a, c, i, j, hs, lim, rmv);
        sf_new(a)[c] = h;
        a.rmv[c] = rmv;
        // computeVolume: depth from the previous step
        const double cellV = p.dx * p.dy * depth_cur(a)[c];
        a.cellV[c] = cellV;
        for (int q = 0; q < p.nScalar; q++)
            a.Sm[(size_t)q * a.sStride + c] = s_cur(a)[(size_t)q * a.sStride + c] * cellV;
        // updateDepth
        const double d = h - a.bottomZ[c];
        depth_nxt(a)[c] = d < 0 ? 0 : d;
    }
    if (FUSED) store_vloss<R, 3>(a, loss, lds);
    else store_vloss<R, 2>(a, lim, lds);
}

// ---------------------------------------------------------------------------
// updateAllFaceDepth + detectWaterfallLocation + updateVelocity +
// waterfallCorrection (nsfunctions.c:143-236, 1080-1166, 250-286) for the
// faces of every Ng index.  A face flag is +1 when the XM/YM test of the cell
// beyond the face holds (that sweep runs last), else -1 when the XP/YP test of
// the cell itself holds.  The rmvCFL zeroing of a cell's own and its west /
// south faces becomes: a face is zeroed when either adjacent interior cell
// has rmvCFL == 1.
template <class R>
__device__ inline double face_depth_x(const StepArgsT<R> &a, int k)
{
    // depthXP per updateAllFaceDepth (irank = 0 = nrank - 1), before the
    // final negative clip
    const Grid &g = a.g;
    const double *h = sf_new(a), *bX = a.bottomXP;
    if (k < g.N) {
        const int i = k % g.nx, j = k / g.nx;
        if (i == g.nx - 1) return 0.0;               // iPbd faces zeroed
        const int e = nbE(g, k, i, j);
        const double hc = h[k], he = h[e];
        return (hc > he ? hc : he) - bX[k];
    }
    int r = k - g.N;
    if (r < 2 * g.nx) {                              // jP / jM ghost rows
        const double hc = h[k], hn = h[k + 1];
        return (hc > hn ? hc : hn) - bX[k];
    }
    return 0.0;                                      // iP / iM ghosts, corners
}

template <class R>
__device__ inline double face_depth_y(const StepArgsT<R> &a, int k)
{
    const Grid &g = a.g;
    const double *h = sf_new(a), *bY = a.bottomYP;
    if (k < g.N) {
        const int i = k % g.nx, j = k / g.nx;
        if (j == g.ny - 1) return 0.0;               // last strip: jPbd zeroed
        const double hc = h[k], hn = h[k + g.nx];
        (void)i;
        return (hc > hn ? hc : hn) - bY[k];
    }
    int r = k - g.N;
    if (r < g.nx) return 0.0;                        // jPgt zeroed (last strip)
    r -= g.nx;
    if (r < g.nx) return 0.0;                        // jMgt zeroed
    r -= g.nx;
    if (r < g.ny) {                                  // iP ghost <- iPbd value
        const int src = r * g.nx + g.nx - 1;
        if (r == g.ny - 1) {
            // copied before the top row is zeroed (nsfunctions.c:214-231)
            const double hc = h[src], hn = h[src + g.nx];
            return (hc > hn ? hc : hn) - bY[src];
        }
        return face_depth_y(a, src);
    }
    r -= g.ny;
    if (r < g.ny) {
        const int src = r * g.nx;
        if (r == g.ny - 1) {
            const double hc = h[src], hn = h[src + g.nx];
            return (hc > hn ? hc : hn) - bY[src];
        }
        return face_depth_y(a, src);
    }
    return 0.0;
}

template <class R>
__device__ inline void faces_cell(const StepArgsT<R> &a, int k, double &loss);

template <class R>
__global__ void __launch_bounds__(256) k_faces(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.nb1) return;
    __shared__ double lds[16];
    if (tail_off(a)) return;
    // blocks from the last index down: the ghost faces, which recompute the
    // faces they mirror, in the first blocks, not in the last ones
    const int blk = a.nb1 - 1 - (int)blockIdx.x;
    const int k = blk * blockDim.x + threadIdx.x;
    double loss[1] = {0.0};
    if (k < a.g.Ng) faces_cell(a, k, loss[0]);
    block_sum<1>(loss, lds);
    if (threadIdx.x == 0) a.vpart[3 * a.maxBlocks + blk] += loss[0];
}

// updateVelocity + waterfallCorrection of interior cell c's east and north
// faces, given their clipped depths; a pure function of the step's inputs,
// so a ghost face can recompute the face it mirrors
struct FaceVel {
    double u, v, fX, fY, loss;
};

template <class R>
__device__ inline FaceVel face_vel(const StepArgsT<R> &a, int c, double dX, double dY)
{
    const Grid &g = a.g;
    const P &p = a.p;
    const int i = c % g.nx, j = c / g.nx;
    const int e = nbE(g, c, i, j), n = nbN(g, c, i, j);
    const double *hs = sf_new(a);
    const double hc = hs[c];
    const double bXc = a.bottomXP[c], bYc = a.bottomYP[c];
    FaceVel r{0, 0, 0, 0, 0};
    // waterfall flags of this cell's east and north faces
    if (hc < bXc - 0.0 && dX >= p.wtfh) r.fX = -1;
    if (e < g.N && hs[e] < bXc - 0.0 && dX >= p.wtfh) r.fX = 1;
    if (hc < bYc - 0.0 && dY >= p.wtfh) r.fY = -1;
    if (n < g.N && hs[n] < bYc - 0.0 && dY >= p.wtfh) r.fY = 1;
    // updateVelocity
    double uc = a.EnXP[c] - (p.g * p.dt / p.dx) * (hs[e] - hc);
    double vc = a.EnYP[c] - (p.g * p.dt / p.dy) * (hs[n] - hc);
    if (dX == 0) uc = 0;
    if (dY == 0) vc = 0;
    const bool rmvC = a.rmv[c] == 1;
    if (rmvC || (e < g.N && a.rmv[e] == 1)) uc = 0;
    if (rmvC || (n < g.N && a.rmv[n] == 1)) vc = 0;
    // waterfallCorrection (Cw = 0.7, CFLmax = 0.24)
    if (r.fX != 0) {
        const double vel0 = uc;
        uc = r.fX * 0.7 * sqrt(2 * p.g * dX);
        if (uc > 0.24 * p.dx / p.dt) uc = 0.24 * p.dx / p.dt;
        r.loss -= (vel0 - uc) * p.dt * dX * p.dy;
    }
    if (r.fY != 0) {
        const double vel0 = vc;
        vc = r.fY * 0.7 * sqrt(2 * p.g * dY);
        if (vc > 0.24 * p.dy / p.dt) vc = 0.24 * p.dy / p.dt;
        r.loss -= (vel0 - vc) * p.dt * dY * p.dx;
    }
    r.u = uc;
    r.v = vc;
    return r;
}

__device__ inline double clipped(double d) { return d < 0 ? 0 : d; }

// one Ng index: its face depths, velocities and flow rates
// (computeFaceFlowRate, nsfunctions.c:350-358).  A ghost face takes the
// velocity of the face it mirrors (enforceVelocityBC, nsfunctions.c:289-314),
// recomputed here from the same inputs, so the flow rates follow in the same
// kernel.  The rmvCFL pass also zeroes the iM-ghost face of a left-column
// cell and the jM-ghost face of a bottom-row cell; enforceVelocityBC
// overwrites those ghosts right after, so nothing is kept of it.
template <class R>
__device__ inline void faces_cell(const StepArgsT<R> &a, int k, double &loss)
{
    const Grid &g = a.g;
    const P &p = a.p;
    const double dX = clipped(face_depth_x(a, k));
    const double dY = clipped(face_depth_y(a, k));
    a.depthXP[k] = dX;
    a.depthYP[k] = dY;
    double u, v;
    if (k >= g.N) {
        // sticky ghost waterfall flags from the XM / YM tests of the adjacent
        // boundary cell (never reset by the reference)
        int r = k - g.N;
        if (r >= g.nx && r < 2 * g.nx) {             // jM ghost: YM face of cell (i, 0)
            const int i = r - g.nx;
            if (sf_new(a)[i] < a.bottomYP[k] - 0.0 && dY >= p.wtfh) a.wtfYP[k] = 1;
        } else if (r >= 2 * g.nx + g.ny && r < 2 * g.nx + 2 * g.ny) {   // iM ghost
            const int jj = r - 2 * g.nx - g.ny;
            if (sf_new(a)[jj * g.nx] < a.bottomXP[k] - 0.0 && dX >= p.wtfh) a.wtfXP[k] = 1;
        }
        const int su = ghost_source(g, k, true), sv = ghost_source(g, k, false);
        u = face_vel(a, su, clipped(face_depth_x(a, su)), clipped(face_depth_y(a, su))).u;
        v = face_vel(a, sv, clipped(face_depth_x(a, sv)), clipped(face_depth_y(a, sv))).v;
    } else {
        const FaceVel f = face_vel(a, k, dX, dY);
        u = f.u;
        v = f.v;
        loss += f.loss;
    }
    a.uuXP[k] = u;
    a.vvYP[k] = v;
    a.Fuu[k] = u * p.dy * dX;
    a.Fvv[k] = v * p.dx * dY;
}

// adjustTidalVelocity (nsfunctions.c:1034-1078) and adjustTidalVelocitySub
// (subgrid.c:1291-1332), folded into the velocity interpolation, the first
// reader of their result.  The YP boundary sets vvYP of each tidal P cell,
// the YM boundary vvYP of the face south of each tidal M cell, from the
// step's fluxes (left as they are: the reference does not recompute them)
// and the cell's surface change.  tidal_cell_v: that value for tidal cell c
// (P: isP), rounded to the state precision as the store rounds it, and
// whether it is set (the plain YM boundary leaves a face alone where the
// flux or the depth is zero).
template <class R, bool SUB>
__device__ inline bool tidal_cell_v(const StepArgsT<R> &a, int c, bool isP, double &v)
{
    const Grid &g = a.g;
    const P &p = a.p;
    const int i = c % g.nx, j = c / g.nx;
    const double Fww = SUB ? (sf_new(a)[c] - sf_start(a)[c]) * a.Z[c] / p.dt
                           : (sf_new(a)[c] - sf_start(a)[c]) * p.dx * p.dy / p.dt;
    const double all = isP ? a.Fvv[nbS(g, c, i, j)] + a.Fuu[nbW(g, c, i, j)] - a.Fuu[c] - Fww
                           : -a.Fvv[c] + a.Fuu[nbW(g, c, i, j)] - a.Fuu[c] - Fww;
    double w;
    if (SUB) {
        const double Oyc = a.Oy[c];
        w = (all != 0 && Oyc > 0) ? (isP ? all / Oyc : -all / Oyc) : 0;
    } else {
        const double dc = depth_nxt(a)[c];
        const bool on = all != 0 && dc > 0;
        if (!isP && !on) return false;
        w = on ? (isP ? all / (dc * p.dx) : -all / (dc * p.dx)) : 0;
    }
    v = static_cast<double>(static_cast<R>(w));
    return true;
}

// the value vvYP[k] holds after adjustTidalVelocity[Sub], for an interior
// face k the velocity interpolation reads: the YP boundary's value where k
// is a tidal P cell (bit 0 of b), then the YM boundary's where the cell
// north of k is a tidal M cell (bit 1; the reference's order), else v, the
// face's value as loaded.  set: k changed.
template <class R, bool SUB>
__device__ inline double tidal_v(const StepArgsT<R> &a, int k, int b, double v, bool &set)
{
    double w;
    if ((b & 1) && tidal_cell_v<R, SUB>(a, k, true, w)) {
        v = w;
        set = true;
    }
    if ((b & 2) && tidal_cell_v<R, SUB>(a, k + a.g.nx, false, w)) {
        v = w;
        set = true;
    }
    return v;
}

// velocityInterp's four vvYP values for cell c (flat-index neighbours, an
// index below 0 reads 0) after the tidal adjustment; the thread stores the
// adjusted values of the faces it owns: vvYP[c], and for a tidal M cell of
// the first row its jM ghost face (no interior thread reads that face).
// Two bytes per cell (tface) say which of the faces are tidal faces and of
// which boundary; the adjustment's stores come after all of its loads.
struct InterpV {
    double c, c1, mn, mn1;
};
template <class R, bool SUB>
__device__ inline InterpV interp_v(const StepArgsT<R> &a, int c, int i, int j)
{
    const Grid &g = a.g;
    const int nx = g.nx;
    const auto v = a.vvYP;
    InterpV r{v[c], v[c + 1], c - nx >= 0 ? v[c - nx] : 0.0, c + 1 - nx >= 0 ? v[c + 1 - nx] : 0.0};
    const int t = a.nbTide > 0 ? a.tface[c] : 0;
    if (t) {
        bool set = false, any = false;
        r.c = tidal_v<R, SUB>(a, c, t, r.c, set);
        r.c1 = tidal_v<R, SUB>(a, c + 1, t >> 2, r.c1, any);
        r.mn = tidal_v<R, SUB>(a, c - nx, t >> 4, r.mn, any);
        r.mn1 = tidal_v<R, SUB>(a, c + 1 - nx, t >> 6, r.mn1, any);
        double w;
        const bool ghost = (t & 256) && tidal_cell_v<R, SUB>(a, c, false, w);
        if (set) v[c] = r.c;
        if (ghost) v[nbS(g, c, i, j)] = w;
    }
    return r;
}

// ---------------------------------------------------------------------------
// volumeByFlux (nsfunctions.c:373-409; volumeByFluxSub subgrid.c:1227-1265,
// whose inflow needs bcType != 3) of cell c, and, with the budgets on, its
// share of the flux-form volume budget: the change of the cell's volume
// minus the flow through its faces on the domain boundary, the inflow and
// the volume the clip of a negative value added.  Interior faces cancel
// between their two cells, so the sum over the domain is zero up to the
// rounding of the cells' updates.
struct VolBudget {
    DD res;
    double vol, bnd, in, clip, moved;
};

template <class R>
__device__ inline void volume_update(const StepArgsT<R> &a, int c, int i, int j, int tt, bool inflowOn, VolBudget &b)
{
    const Grid &g = a.g;
    const P &p = a.p;
    const int w = nbW(g, c, i, j), s = nbS(g, c, i, j);
    const double Vold = a.cellV[c];
    double V = Vold + p.dt * (a.Fuu[w] - a.Fuu[c] + a.Fvv[s] - a.Fvv[c]);
    const int ninflow = inflowOn ? (a.flags[c] >> 2) : 0;
    for (int q = 0; q < ninflow; q++) V = V + a.inflow[tt] * p.dt;
    double clip = 0.0;
    if (V < 0) {
        clip = -V;
        V = 0;
    }
    a.cellV[c] = V;
    if (p.budget) {
        double bf = 0.0;
        if (i == 0) bf += a.Fuu[w];
        if (i == g.nx - 1) bf -= a.Fuu[c];
        if (j == 0) bf += a.Fvv[s];
        if (j == g.ny - 1) bf -= a.Fvv[c];
        const double bnd = p.dt * bf, in = ninflow * (a.inflow[tt] * p.dt);
        DD r = dd_two_sum(V, -Vold);
        r = dd_add(r, -bnd);
        r = dd_add(r, -in);
        r = dd_add(r, -clip);
        b.res = dd_add(b.res, r);
        b.vol += V;
        b.bnd += bnd;
        b.in += in;
        b.clip += clip;
        b.moved += fabs(V - Vold);
    }
}

// the block's volume-budget partials (budget slots 0..6)
template <class R>
__device__ inline void store_vol_budget(const StepArgsT<R> &a, VolBudget &b, double *lds, int slot)
{
    const DD r = block_dd_sum(b.res, lds);
    double v[5] = {b.vol, b.bnd, b.in, b.clip, b.moved};
    block_sum<5>(v, lds);
    if (threadIdx.x == 0) {
        const int mb = a.maxBlocks, k = slot;
        a.bpart[0 * mb + k] = r.hi;
        a.bpart[1 * mb + k] = r.lo;
        for (int q = 0; q < 5; q++) a.bpart[(2 + q) * mb + k] = v[q];
    }
}

// ---------------------------------------------------------------------------
// velocityInterp + volumeByFlux + updateCD + thinLayerDrag + monitorCFL
// (nsfunctions.c:337-347, 373-409, 1170-1189, 691-727, 1192-1232).
// velocityInterp uses flat-index neighbours (row wrap, ghost reads); indices
// below 0 read 0.
template <class R>
__global__ void __launch_bounds__(256) k_finish(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.nb1) return;
    __shared__ double lds[16 * 5];
    if (tail_off(a)) return;
    const Grid g = a.g;
    // blocks start at the tidal boundary (row_block), whose faces take
    // extra work here
    const int blk = a.rowsDown ? a.nb1 - 1 - (int)blockIdx.x : (int)blockIdx.x;
    const int c = blk * blockDim.x + threadIdx.x;
    const P &p = a.p;
    double cflx = 0, cfly = 0;
    VolBudget vb{{0.0, 0.0}, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (c < g.N) {
        const int tt = cur_tt(a);
        const int nx = g.nx, i = c % nx, j = c / nx;
        const auto u = a.uuXP;
        const double um1 = c - 1 >= 0 ? u[c - 1] : 0.0;
        const double um1n = c - 1 + nx >= 0 ? u[c - 1 + nx] : 0.0;
        const InterpV v = interp_v<R, false>(a, c, i, j);
        a.uuYP[c] = 0.25 * (u[c] + um1 + u[c + nx] + um1n);
        a.vvXP[c] = 0.25 * (v.c + v.c1 + v.mn + v.mn1);

        volume_update(a, c, i, j, tt, true, vb);

        const double d = depth_nxt(a)[c];
        if (p.CDnotN == 0 && d > 0) {
            const double n = a.manning ? a.manning[c] : p.manningN;
            const double coeff = p.g * n * n;
            const double cd = coeff / pow(d, 1.0 / 3.0);
            a.CDXP[c] = cd;
            a.CDYP[c] = cd;
        }
        if (p.useThinLayer == 1 && p.CDnotN == 0 && d > 0) {
            const double n = a.manning ? a.manning[c] : p.manningN;
            const double coeff = p.g * n * n;
            if (a.depthXP[c] <= p.hD) a.CDXP[c] = coeff / pow(d, 2.0 / 3.0);
            if (a.depthYP[c] <= p.hD) a.CDYP[c] = coeff / pow(d, 2.0 / 3.0);
        }
        cflx = fabs(u[c]) * p.dt / p.dx;
        cfly = fabs(v.c) * p.dt / p.dy;
        a.CFLx[c] = cflx;
        a.CFLy[c] = cfly;
    }
    // thinLayerDrag with constant CD covers every Ng face
    if (p.useThinLayer == 1 && p.CDnotN == 1 && c < g.Ng) {
        if (a.depthXP[c] <= p.hD) a.CDXP[c] = p.CDx + (p.CDmax - p.CDx) * (p.hD - a.depthXP[c]) / p.hD;
        if (a.depthYP[c] <= p.hD) a.CDYP[c] = p.CDy + (p.CDmax - p.CDy) * (p.hD - a.depthYP[c]) / p.hD;
    }
    const double mx = block_max(cflx, lds);
    const double my = block_max(cfly, lds);
    if (threadIdx.x == 0) {
        a.cflpart[blk] = mx;
        a.cflpart[a.maxBlocks + blk] = my;
        cfl_stop_test(a, mx, my);
    }
    if (p.budget) store_vol_budget(a, vb, lds, blk);
}

// ---------------------------------------------------------------------------
// Scalar transport (scalar.c:30-56, upwind or TVD): scalarAdvection +
// scalarDiffusion + updateScalar's connectivity limiter, inflow mass and
// concentration, and the out-of-range reset (scalar.c:68-97, 259-416), then
// enforceScalarBC (scalar.c:419-471).  Neighbours read the step's start
// buffer, the new concentration goes to the other one: the interior with
// the tidal Dirichlet values and zero where dry, and the ghosts mirroring a
// boundary cell (corners untouched) with that cell's value before those.
// superbee limiter of the TVD scheme (scalar.c:113-117)
__device__ inline double superbee(double r)
{
    double p1, p2, phi;
    if (2 * r > 1) p1 = 1; else p1 = 2 * r;
    if (r > 2) p2 = 2; else p2 = r;
    if (p1 > 0) phi = p1; else phi = 0;
    if (p2 > phi) phi = p2;
    return phi;
}

// face value of the TVD scheme (scalar.c:103-245) for the face between cell
// L and the next cell H (x or y): F the flow rate at the face, FL the one at
// the face on L's far side, FH at H's far side; SLL, SHH the cells one
// further out; Cu the face's Courant number.  Upwind from L when F > 0
// (first order if the flow at L's far side is not also positive), else from H.
__device__ inline double tvd_face(double SL, double SH, double SLL, double SHH, double F, double FL, double FH,
                                  double Cu)
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

// One scalar's share of the flux-form mass budget of a cell: the change of
// its mass minus the advective and diffusive fluxes through its faces on the
// domain boundary and the inflow mass (the interior faces cancel between
// their two cells).
struct MassBudget {
    DD res;
    double mass, bnd, in, moved;
};

template <class R, bool SUB, bool BUDGET>
__device__ inline void scalar_body(const StepArgsT<R> &a, double *lds)
{
    if (tail_off(a) || a.p.nScalar < 1) return;
    const Grid g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = row_block(a) * blockDim.y + threadIdx.y;
    // every thread of the block stays to the end (the budget's block sums);
    // the cells outside the grid do nothing
    const bool valid = i < g.nx && j < g.ny;
    const P &p = a.p;
    const int tt = cur_tt(a);
    const int c = valid ? j * g.nx + i : 0;
    const int e = valid ? nbE(g, c, i, j) : 0, w = valid ? nbW(g, c, i, j) : 0;
    const int n = valid ? nbN(g, c, i, j) : 0, s = valid ? nbS(g, c, i, j) : 0;
    const auto Fu = a.Fuu;
    const auto Fv = a.Fvv;
    const double dt = p.dt, dx = p.dx, dy = p.dy;
    // geometry of the cell, the same for every scalar: flow rates, face
    // areas or depths, the connectivity with the previous-step surface
    // (scalar.c:292-341; the subgrid tests face areas, subgrid.c:1588-1630)
    const double *ho = sf_start(a), *zb = a.bottomZ;
    const auto fx = SUB ? a.Nx : a.depthXP;
    const auto fy = SUB ? a.Oy : a.depthYP;
    double Fuc = 0, Fuw = 0, Fvc = 0, Fvs = 0, fxc = 0, fxw = 0, fyc = 0, fys = 0, V = 0;
    bool cE = false, cW = false, cN = false, cS = false;
    int ninflow = 0;
    uint8_t fl = 0;
    if (valid) {
        Fuc = Fu[c]; Fuw = Fu[w]; Fvc = Fv[c]; Fvs = Fv[s];
        fxc = fx[c]; fxw = fx[w]; fyc = fy[c]; fys = fy[s];
        cE = fxc > 0 && ho[e] - zb[e] > 0;
        cW = fxw > 0 && ho[w] - zb[w] > 0;
        cN = n <= g.N && fyc > 0 && ho[n] - zb[n] > 0;
        cS = s <= g.N && fys > 0 && ho[s] - zb[s] > 0;
        // the subgrid path adds inflow only when bcType != 3 (subgrid.c:1660)
        ninflow = (SUB && p.bcType == 3) ? 0 : (a.flags[c] >> 2);
        V = a.cellV[c];
        fl = a.flags[c];
    }
    const bool bottomRow = (j == 0), topRowNotFirst = (j == g.ny - 1 && i > 0);
    // each scalar in turn, as separate passes would run them: a pass that
    // resets an out-of-range cell lowers that cell's surface and depth for
    // the passes after it
    for (int q = 0; q < p.nScalar; q++) {
        MassBudget mb{{0.0, 0.0}, 0.0, 0.0, 0.0, 0.0};
        if (valid) {
        const auto S = s_cur(a) + (size_t)q * a.sStride;
        double *Sm = a.Sm + (size_t)q * a.sStride;
        const double Sc = S[c], Se = S[e], Sw = S[w], Sn = S[n], Ss = S[s];
        const double smOld = Sm[c];
        double sm = smOld;
        double Sd, Su, Sr, Sl;
        {
            if (p.scalarAdv == 2) {
                // TVD superbee (scalar.c:99-255)
                const int ee = nbEE(g, c, i, j), ww = nbWW(g, c, i, j), nn = nbNN(g, c, i, j), ss = nbSS(g, c, i, j);
                const auto u = a.uuXP;
                const auto v = a.vvYP;
                Sd = tvd_face(Sc, Se, Sw, S[ee], Fuc, Fuw, Fu[e], u[c] * dt / dx);
                Su = tvd_face(Sw, Sc, S[ww], Se, Fuw, Fu[ww], Fuc, u[w] * dt / dx);
                Sr = tvd_face(Sc, Sn, Ss, S[nn], Fvc, Fvs, Fv[n], v[c] * dt / dy);
                // the YM face's Courant number divides by dx (scalar.c:222)
                Sl = tvd_face(Ss, Sc, S[ss], Sn, Fvs, Fv[ss], Fvc, v[s] * dt / dx);
            } else {
                Sd = Fuc > 0 ? Sc : Se;
                Su = Fuw > 0 ? Sw : Sc;
                Sr = Fvc > 0 ? Sc : Sn;
                Sl = Fvs > 0 ? Ss : Sc;
            }
            sm = sm + dt * (-Fuc * Sd + Fuw * Su - Fvc * Sr + Fvs * Sl);
        }
        // diffusion coefficients of the four faces
        const double kE = SUB ? p.Kx * fxc / dx : p.Kx * dy * fxc / dx;
        const double kW = SUB ? p.Kx * fxw / dx : p.Kx * dy * fxw / dx;
        const double kN = SUB ? p.Ky * fyc / dy : p.Ky * dx * fyc / dy;
        const double kS = SUB ? p.Ky * fys / dy : p.Ky * dx * fys / dy;
        // scalarDiffusionSub, subgrid.c:1552-1567 / scalarDiffusion, scalar.c:259-274
        sm = sm + dt * (kE * (Se - Sc) - kW * (Sc - Sw) + kN * (Sn - Sc) - kS * (Sc - Ss));
        double vals[4];
        int m = 0;
        if (cE) vals[m++] = Se;
        if (cW) vals[m++] = Sw;
        if (cN) vals[m++] = Sn;
        if (cS) vals[m++] = Ss;
        double smax = 0, smin = 0;
        if (m) {
            smax = vals[0];
            smin = vals[0];
            for (int t = 1; t < m; t++) {
                if (vals[t] > smax) smax = vals[t];
                if (vals[t] < smin) smin = vals[t];
            }
        }
        for (int t = 0; t < ninflow; t++) sm = sm + p.inflowS * a.inflow[tt] * dt / p.inflowLocLength;
        Sm[c] = sm;
        if (BUDGET && p.budget) {
            double bf = 0.0;
            if (i == 0) bf += Fuw * Su - kW * (Sc - Sw);
            if (i == g.nx - 1) bf += -Fuc * Sd + kE * (Se - Sc);
            if (j == 0) bf += Fvs * Sl - kS * (Sc - Ss);
            if (j == g.ny - 1) bf += -Fvc * Sr + kN * (Sn - Sc);
            const double bnd = dt * bf;
            const double in = ninflow ? ninflow * (p.inflowS * a.inflow[tt] * dt / p.inflowLocLength) : 0.0;
            DD r = dd_two_sum(sm, -smOld);
            r = dd_add(r, -bnd);
            r = dd_add(r, -in);
            mb.res = r;
            mb.mass = sm;
            mb.bnd = bnd;
            mb.in = in;
            mb.moved = fabs(sm - smOld);
        }
        double Sv = V > 0 ? sm / V : 0;
        if (m != 0 && Sv > smax && ninflow == 0) Sv = smax;
        else if (m != 0 && Sv < smin && ninflow == 0) Sv = smin;
        if (Sv > 100 || Sv < 0) {
            if (!bottomRow && !topRowNotFirst) {
                Sv = 0;
                sf_new(a)[c] = sf_new(a)[c] - depth_nxt(a)[c];
                depth_nxt(a)[c] = 0;
            }
        }
        const auto Snx = s_nxt(a) + (size_t)q * a.sStride;
        if (j == g.ny - 1) Snx[gjP(g, i)] = Sv;
        if (j == 0) Snx[gjM(g, i)] = Sv;
        if (i == g.nx - 1) Snx[giP(g, j)] = Sv;
        if (i == 0) Snx[giM(g, j)] = Sv;
        if ((fl & 1) && p.bcType != 2) Sv = p.useConstTidePS == 1 ? p.tidePS : a.tidalPS[tt];
        if ((fl & 2) && p.bcType != 1) Sv = p.useConstTideMS == 1 ? p.tideMS : a.tidalMS[tt];
        if (depth_nxt(a)[c] <= 0) Sv = 0;
        Snx[c] = Sv;
        }
        if (BUDGET && p.budget) {
            // the block's partials of this scalar (kBudgetVolSlots + kBudgetScalarSlots q ...)
            const DD r = block_dd_sum(mb.res, lds);
            double v[4] = {mb.mass, mb.bnd, mb.in, mb.moved};
            block_sum<4>(v, lds);
            if (threadIdx.x == 0 && threadIdx.y == 0) {
                const int mbk = a.maxBlocks, k = blockIdx.x + a.gx2 * row_block(a);
                const int s0 = kBudgetVolSlots + kBudgetScalarSlots * q;
                a.bpart[(size_t)(s0 + 0) * mbk + k] = r.hi;
                a.bpart[(size_t)(s0 + 1) * mbk + k] = r.lo;
                for (int u = 0; u < 4; u++) a.bpart[(size_t)(s0 + 2 + u) * mbk + k] = v[u];
            }
        }
    }
}

// every scalar of a member in one launch (the geometry is read once)
// (the mass budget's block sums compiled only into the variant a group with
// a budgeting member launches)
template <class R, bool BUDGET>
__global__ void __launch_bounds__(256) k_scalar(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.gx2 || blockIdx.y >= a.gy2) return;
    __shared__ double lds[16 * 4];
    scalar_body<R, false, BUDGET>(a, lds);
}
template <class R, bool BUDGET>
__global__ void __launch_bounds__(256) k_scalar_sub(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.gx2 || blockIdx.y >= a.gy2) return;
    __shared__ double lds[16 * 4];
    scalar_body<R, true, BUDGET>(a, lds);
}

// the conservation budgets of the step: the blocks' partials summed in
// block order (double-double for the residuals), the step's residuals and
// totals added to the member's accumulators (thread 0)
template <class R>
__device__ inline void step_end_budget(const StepArgsT<R> &a)
{
    __shared__ double lds[16 * 5];
    const int t = threadIdx.x, nt = blockDim.x, mb = a.maxBlocks;
    Budget &bu = a.ctl->budget;
    {
        DD r{0.0, 0.0};
        double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int b = t; b < a.nb1; b += nt) {
            r = dd_add(r, DD{a.bpart[b], a.bpart[mb + b]});
            for (int q = 0; q < 5; q++) v[q] += a.bpart[(size_t)(2 + q) * mb + b];
        }
        r = block_dd_sum(r, lds);
        block_sum<5>(v, lds);
        if (t == 0) {
            const double res = r.hi + r.lo;
            bu.vol = v[0];
            bu.volRes += res;
            bu.volResMax = fmax(bu.volResMax, fabs(res));
            bu.volBnd += v[1];
            bu.volIn += v[2];
            bu.volClip += v[3];
            bu.volMoved += v[4];
            bu.steps += 1;
        }
    }
    for (int q = 0; q < a.p.nScalar; q++) {
        const int s0 = kBudgetVolSlots + kBudgetScalarSlots * q;
        DD r{0.0, 0.0};
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        for (int b = t; b < a.nb2; b += nt) {
            r = dd_add(r, DD{a.bpart[(size_t)s0 * mb + b], a.bpart[(size_t)(s0 + 1) * mb + b]});
            for (int u = 0; u < 4; u++) v[u] += a.bpart[(size_t)(s0 + 2 + u) * mb + b];
        }
        r = block_dd_sum(r, lds);
        block_sum<4>(v, lds);
        if (t == 0) {
            const double res = r.hi + r.lo;
            bu.mass[q] = v[0];
            bu.massRes[q] += res;
            bu.massResMax[q] = fmax(bu.massResMax[q], fabs(res));
            bu.massBnd[q] += v[1];
            bu.massIn[q] += v[2];
            bu.massMoved[q] += v[3];
        }
    }
}

// A step entered at its update phase from a state given at a step boundary
// (a parity check starting from a mid-step state): the given surface is the
// solve's new one, so the two surface buffers trade places (sf_new takes the
// new surface, sf_start the old).  (Plain pointers: a kernel parameter typed
// for the global address space would be named differently in the device
// pass than in the host pass, and the launch would not find the kernel.)
__global__ void __launch_bounds__(256) k_swap_surface(double *a, double *b, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double t = a[k];
    a[k] = b[k];
    b[k] = t;
}

// Step end with the conservation budgets on: the step's budget partials
// reduced in block order and added to the member's accumulators.  (The
// volume losses accumulate in per-block slots, the CFL stop test runs in
// the velocity update and the step's bookkeeping in the flag gather and
// the next step's begin, so a step without budgets has no such launch.)
// One block of 1024 threads.
template <class R>
__global__ void __launch_bounds__(1024) k_step_end(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (tail_off(a) || !a.p.budget) return;
    step_end_budget(a);
}

}  // namespace frehd

// ===========================================================================
// Subgrid bathymetry (useSubgrid = 1, staggeredV = 1, useminA = 0,
// useSubDrag = 0): the step kernels that differ from the plain grid.
// ===========================================================================
namespace frehd {

// interpSurf, utilities.c:155-162
__device__ inline double interp_surf(double x, double x1, double y1, double x2, double y2)
{
    if (x2 == x1) return y1;
    return (y2 - y1) * (x - x2) / (x2 - x1) + y2;
}

// The inputs of cell c's lookup that do not depend on its new surface,
// loaded together before the surface is known: the lookup's chain of
// dependent memory round trips is then this one, the level walk (cached)
// and the table reads.  edgeX/edgeY of a ghost neighbour read 0 (the
// reference reads past its N-array); the face bottoms of a ghost neighbour
// are not read.
struct SubPre {
    double hSub, eXc, eXw, eYc, eYs, bXc, bYc, bXw, bYs;
    int ind;
    uint8_t dSub;
};

template <class R>
__device__ inline SubPre subarea_pre(const StepArgsT<R> &a, int c, int i, int j)
{
    const Grid &g = a.g;
    const int N = g.N;
    const int w = nbW(g, c, i, j), s = nbS(g, c, i, j);
    SubPre q;
    q.hSub = a.hSub[c];
    q.dSub = a.dSub[c];
    q.ind = a.ind[c];
    q.eXc = a.edgeX[c];
    q.eXw = w < N ? a.edgeX[w] : 0.0;
    q.eYc = a.edgeY[c];
    q.eYs = s < N ? a.edgeY[s] : 0.0;
    q.bXc = a.bottomXP[c];
    q.bYc = a.bottomYP[c];
    q.bXw = w < N ? a.bottomXP[w] : 0.0;
    q.bYs = s < N ? a.bottomYP[s] : 0.0;
    return q;
}

// computeSubArea for cell c (subgrid.c:55-251): walk the level index to the
// surface, interpolate the half-cell volumes and face areas between the two
// bracketing levels, zero what is dry.
template <class R, bool OTF>
__device__ inline void subarea_cell(const StepArgsT<R> &a, int c, int i, int j, double h, double depthOld,
                                    const SubPre &q)
{
    // the outputs are a function of (h, depthOld == 0) and constant inputs,
    // and the level walk from the last index with the same h stops at once:
    // bitwise-equal inputs give bitwise-equal outputs, so skip the lookup
    const uint8_t dry = depthOld == 0;
    if (__double_as_longlong(q.hSub) == __double_as_longlong(h) && q.dSub == dry) return;
    a.hSub[c] = h;
    a.dSub[c] = dry;
    const Grid &g = a.g;
    const P &p = a.p;
    const int N = g.N, Ne = p.Ne;
    const double dsurf = p.dsurf;
    const double *as = a.allSurf;
    int ind = q.ind;
    while (fabs(as[ind] - h) > 0.5 * dsurf + dsurf / 100) {
        if (h > as[ind]) ind += 1;
        else ind -= 1;
        if (ind < 0) { ind = 0; break; }
        if (ind > Ne - 1) { ind = Ne - 1; break; }
    }
    a.ind[c] = ind;
    const int j1 = ind;
    const int j2 = j1 == 0 ? 0 : (j1 == Ne - 1 ? Ne - 1 : (h > as[j1] ? j1 + 1 : j1 - 1));
    // drag-model values at the lower level (subgrid.c:121-127)
    if (p.useSubDrag == 2) {
        a.YhX[c] = a.allYhX[(size_t)j1 * N + c];
        a.YhY[c] = a.allYhY[(size_t)j1 * N + c];
    } else if (p.useSubDrag == 3 || p.useSubDrag == 4) {
        a.Cn[c] = a.allCn[(size_t)j1 * N + c];
    }
    const double s1 = as[j1], s2 = as[j2];
    // the two levels' 10 values each: adjacent in the interleaved tables, or
    // evaluated from the cell's pixels
    // (tables in the state precision: evaluated values are rounded to it,
    // so both hold the same numbers)
    double t1[10], t2[10];
    if (!OTF || a.tabs) {
        const auto p1 = a.tabs + tab_index(c, j1, N, 0), p2 = a.tabs + tab_index(c, j2, N, 0);
#pragma unroll
        for (int q = 0; q < 10; q++) {
            t1[q] = p1[q];
            t2[q] = p2[q];
        }
    } else {
        const double hl[2] = {level_raw(p.surfmin, dsurf, j1), level_raw(p.surfmin, dsurf, j2)};
        double v[2][10];
        switch (a.bed.r) {            // the common pixel counts unrolled
        case 4: cell_levels<2, 4>(a.bed, i, j, hl, v); break;
        default: cell_levels<2>(a.bed, i, j, hl, v); break;
        }
#pragma unroll
        for (int q = 0; q < 10; q++) {
            t1[q] = (R)v[0][q];
            t2[q] = (R)v[1][q];
        }
    }
    a.Vraw[c] = interp_surf(h, s1, t1[T_V], s2, t2[T_V]);
    a.Zraw[c] = interp_surf(h, s1, t1[T_Z], s2, t2[T_Z]);
    double vxp = interp_surf(h, s1, t1[T_VXP], s2, t2[T_VXP]);
    double vyp = interp_surf(h, s1, t1[T_VYP], s2, t2[T_VYP]);
    double vxm = interp_surf(h, s1, t1[T_VXM], s2, t2[T_VXM]);
    double vym = interp_surf(h, s1, t1[T_VYM], s2, t2[T_VYM]);
    if (p.minTabs) {
        // minimum half-cell volumes and minimum / maximum face areas
        // (subgrid.c:143-150, 210-220), level tables of their own
        const size_t k1 = (size_t)j1 * N + c, k2 = (size_t)j2 * N + c;
        if (p.useminV != 0) {
            a.minVxp[c] = interp_surf(h, s1, a.allMinVxp[k1], s2, a.allMinVxp[k2]);
            a.minVyp[c] = interp_surf(h, s1, a.allMinVyp[k1], s2, a.allMinVyp[k2]);
            a.minVxm[c] = interp_surf(h, s1, a.allMinVxm[k1], s2, a.allMinVxm[k2]);
            a.minVym[c] = interp_surf(h, s1, a.allMinVym[k1], s2, a.allMinVym[k2]);
        }
        if (p.useminA != 0) {
            a.Nmin[c] = interp_surf(h, s1, a.allNmin[k1], s2, a.allNmin[k2]);
            a.Omin[c] = interp_surf(h, s1, a.allOmin[k1], s2, a.allOmin[k2]);
            a.Nmax[c] = interp_surf(h, s1, a.allNmax[k1], s2, a.allNmax[k2]);
            a.Omax[c] = interp_surf(h, s1, a.allOmax[k1], s2, a.allOmax[k2]);
        }
    }
    const int w = nbW(g, c, i, j), s = nbS(g, c, i, j);
    const double wd[4] = {q.eXc, q.eXw, q.eYc, q.eYs};
    const int tid[4] = {T_NP, T_NM, T_OP, T_OM};
    double f[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (wd[q] > s1 && wd[q] < s2) {
            f[q] = wd[q] > h ? 0.0 : interp_surf(h, wd[q], 0.0, s2, t2[tid[q]]);
        } else {
            f[q] = interp_surf(h, s1, t1[tid[q]], s2, t2[tid[q]]);
        }
    }
    if (depthOld == 0) { vxp = 0; vyp = 0; vxm = 0; vym = 0; }
    if (h - q.bXc <= 0) f[0] = 0;
    if (h - q.bYc <= 0) f[2] = 0;
    if (w < N && h - q.bXw <= 0) f[1] = 0;
    if (s < N && h - q.bYs <= 0) f[3] = 0;
    a.Vxp[c] = vxp; a.Vyp[c] = vyp; a.Vxm[c] = vxm; a.Vym[c] = vym;
    a.Np[c] = f[0]; a.Nm[c] = f[1]; a.Op[c] = f[2]; a.Om[c] = f[3];
}

// Init: computeSubArea on the initial surface (initialize.c:107)
template <class R>
__global__ void __launch_bounds__(256) k_subarea_init(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.nbN) return;
    const Grid g = a.g;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const int i = c % g.nx, j = c / g.nx;
    subarea_cell<R, true>(a, c, i, j, sf_start(a)[c], depth_cur(a)[c], subarea_pre(a, c, i, j));
}

// limiterCFL + computeVolumeSub + computeSubArea + updateDepth
// (nsfunctions.c:452-495, subgrid.c:1215-1224, 55-251, nsfunctions.c:239-247)
// FUSED: the surface update is made here too (see k_limiter); OTF: a
// member may evaluate its lookups on the fly from the pixels (the variant
// for tables only keeps the registers, and the occupancy, of that path)
template <class R, bool FUSED, bool OTF>
__global__ void __launch_bounds__(256) k_limiter_sub(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.gx2 || blockIdx.y >= a.gy2) return;
    __shared__ double lds[16 * 3];
    if (tail_off(a)) return;
    // the new depth goes to the second buffer: the next step starts from it
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) a.ctl->flipPending = 1;
    const Grid g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = row_block(a) * blockDim.y + threadIdx.y;
    double loss[3] = {0.0, 0.0, 0.0};
    double (&lim)[2] = *reinterpret_cast<double (*)[2]>(loss);
    if (i < g.nx && j < g.ny) {
        const P &p = a.p;
        const int c = j * g.nx + i;
        // every load that does not depend on the new surface first, so that
        // one round trip covers them (a wave waits at the first use of each
        // load, and a load after a store to another array cannot be moved
        // above it)
        const SubPre pre = subarea_pre(a, c, i, j);
        const double cellV = a.V[c];   // the previous step's subgrid volume
        const double zb = a.bottomZ[c];
        const double dOld = depth_cur(a)[c];
        constexpr int NS = 4;          // scalars loaded up front
        const size_t ss = a.sStride;
        double sv[NS];
#pragma unroll
        for (int q = 0; q < NS; q++) sv[q] = q < p.nScalar ? s_cur(a)[q * ss + c] : 0.0;
        uint8_t rmv;
        const double hs = FUSED ? fused_surface(a, c, i, j, loss[2]) : sf_new(a)[c];
        const double h = limiter_cell(a, c, i, j, hs, lim, rmv);
        const double d = h - zb;
        sf_new(a)[c] = h;
        a.rmv[c] = rmv;
        a.cellV[c] = cellV;
#pragma unroll
        for (int q = 0; q < NS; q++)
            if (q < p.nScalar) a.Sm[q * ss + c] = sv[q] * cellV;
        for (int q = NS; q < p.nScalar; q++) a.Sm[q * ss + c] = s_cur(a)[q * ss + c] * cellV;
        depth_nxt(a)[c] = d < 0 ? 0 : d;
        subarea_cell<R, OTF>(a, c, i, j, h, dOld, pre);
    }
    if (FUSED) store_vloss<R, 3>(a, loss, lds);
    else store_vloss<R, 2>(a, lim, lds);
}

// combineSubArea (subgrid.c:254-462 and the clip loop :466-478) of interior
// cell c, plus the interior faces the boundary loop closes: Nx of the right
// column and Oy of the non-tidal top row.  useminV = 1 changes volumes only
// where Nmin > 0, which stays 0 with useminA = 0 (min_adjust runs when the
// minimum tables are on the device).  A pure function of computeSubArea's
// outputs, so a ghost can recompute the cell it mirrors.
struct Combined {
    double Nx, Oy, Vx, Vy, Zx, Zy, V, Z;
    double OyOpen;            // Oy before the top row's closure
};

// effective-volume factor of combineSubArea (subgrid.c:377-403): r the
// ratio of the minimum to the maximum face area
__device__ inline double volume_lambda(double r, double cap)
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

// useminA / useminV adjustments of combineSubArea (subgrid.c:360-414), after
// the waterfall halving and before the small-value clip; the reference's
// local beta is 0, so useminA takes its first branch (pull Nx, Oy above
// Nmax, Omax back by Config.beta)
template <class R>
__device__ inline void min_adjust(const StepArgsT<R> &a, int c, int e, int n, double &Nx, double &Oy, double &Vx,
                                  double &Vy)
{
    const P &p = a.p;
    const int N = a.g.N;
    const double epsi = 0.0001;
    double minVx = 0.0, minVy = 0.0;
    if (p.useminV != 0) {
        minVx = e < N ? a.minVxp[c] + a.minVxm[e] : 2 * a.minVxp[c];
        minVy = n < N ? a.minVyp[c] + a.minVym[n] : 2 * a.minVyp[c];
    }
    double Nmin = 0.0, Nmax = 0.0, Omin = 0.0, Omax = 0.0;
    if (p.useminA == 1) {
        Nmin = a.Nmin[c]; Nmax = a.Nmax[c]; Omin = a.Omin[c]; Omax = a.Omax[c];
        if (Nx - Nmax > epsi && Nmax > 0) Nx = Nx - p.beta * (Nx - Nmax);
        if (Oy - Omax > epsi && Omax > 0) Oy = Oy - p.beta * (Oy - Omax);
    }
    if (p.useminV == 1) {
        const double Vfloor = 1.0;
        if (Nmin > 0 && Vx > minVx && minVx > 0) {
            const double r = Nmin / Nmax;
            const double lambda = volume_lambda(r, 1.35);
            Vx = Vx - (1 - r) * lambda * Vx;
            if (Vx < Vfloor) Vx = Vfloor;
        }
        if (Omin > 0 && Vy > minVy && minVy > 0) {
            const double r = Omin / Omax;
            const double lambda = volume_lambda(r, 1.4);
            Vy = Vy - (1 - r) * lambda * Vy;
            if (Vy < Vfloor) Vy = Vfloor;
        }
    } else if (p.useminV == 2) {
        Vx = minVx;
        Vy = minVy;
    }
}

// (hs: the surface the lookups were made at -- sf_start at Init, sf_new in
// the step)
template <class R>
__device__ inline Combined combine_cell(const StepArgsT<R> &a, int c, gptr<const double> hs)
{
    const Grid &g = a.g;
    const P &p = a.p;
    const int i = c % g.nx, j = c / g.nx, N = g.N;
    const int e = nbE(g, c, i, j), n = nbN(g, c, i, j);
    const double dxdy = p.dx * p.dy;
    const double Zc = a.Zraw[c], h = hs[c];
    double Nx, Oy, Vx, Vy, Zx, Zy;
    if (e < N) { const double A = a.Np[c], B = a.Nm[e]; Nx = A >= B ? A : B; }
    else Nx = a.Np[c];
    if (n < N) { const double A = a.Op[c], B = a.Om[n]; Oy = A >= B ? A : B; }
    else Oy = a.Op[c];
    if (e < N) {
        const double vp = a.Vxp[c], vm = a.Vxm[e];
        if (vp > 0 && vm > 0) Vx = vp + vm;
        else Vx = 2.0 * (vp + vm);
        Zx = Zc;
        if (Vx > 0 && Zx <= 0) Zx = dxdy;
        if (h < a.bottomXP[c]) { Vx = vm; Zx = 0.5 * Zc; }
        if (hs[e] < a.bottomXP[c]) { Vx = vp; Zx = 0.5 * Zc; }
    } else {
        Vx = 2 * a.Vxp[c];
        Zx = 2 * Zc;
    }
    if (n < N) {
        const double vp = a.Vyp[c], vm = a.Vym[n];
        if (vp > 0 && vm > 0) Vy = vp + vm;
        else Vy = 2.0 * (vp + vm);
        Zy = Zc;
        if (Vy > 0 && Zy <= 0) Zy = dxdy;
        if (h < a.bottomYP[c]) { Vy = vm; Zy = 0.5 * Zc; }
        if (hs[n] < a.bottomYP[c]) { Vy = vp; Zy = 0.5 * Zc; }
    } else {
        Vy = 2 * a.Vyp[c];
        Zy = 2 * Zc;
    }
    if (p.minTabs) min_adjust(a, c, e, n, Nx, Oy, Vx, Vy);
    const double Vmin = 0.002, Amin = 0.002, Azmin = 0.002;
    double V = a.Vraw[c], Z = Zc;
    if (Vx < Vmin) Vx = 0.0;
    if (Vy < Vmin) Vy = 0.0;
    if (V < Vmin) V = 0.0;
    if (Z < Azmin) Z = 1.0;
    if (Zx < Amin) Zx = 0.0;
    if (Zy < Amin) Zy = 0.0;
    if (Nx < Amin) Nx = 0.0;
    if (Oy < Amin) Oy = 0.0;
    const double OyOpen = Oy;
    if (i == g.nx - 1) Nx = 0;                                    // subgrid.c:486
    if (j == g.ny - 1 && p.bcType != 2 && !(a.flags[c] & 1)) Oy = 0;   // :508-524
    return Combined{Nx, Oy, Vx, Vy, Zx, Zy, V, Z, OyOpen};
}

// combineSubArea over the interior (blocks [0, nbN)) and its ghost copies
// (subgrid.c:480-506) over the 2nx + 2ny ghosts (the next nbGhost blocks),
// each ghost recomputing the cell it mirrors.  The iM/iP ghosts of the top
// row take the corner cell's Oy before the top-row closure (the closure
// runs after the ghost copies, subgrid.c:508-524).
template <class R>
__global__ void __launch_bounds__(256) k_combine(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.nbN + a.nbGhost) return;
    const Grid g = a.g;
    if ((int)blockIdx.x < a.nbN) {
        const int c = blockIdx.x * blockDim.x + threadIdx.x;
        if (c >= g.N) return;
        const Combined v = combine_cell(a, c, sf_start(a));
        a.Nx[c] = v.Nx; a.Oy[c] = v.Oy; a.Vx[c] = v.Vx; a.Vy[c] = v.Vy;
        a.Zx[c] = v.Zx; a.Zy[c] = v.Zy; a.V[c] = v.V; a.Z[c] = v.Z;
        return;
    }
    const int k = g.N + (blockIdx.x - a.nbN) * blockDim.x + threadIdx.x;
    if (k >= g.Ng - 4) return;
    const int src = ghost_source(g, k, false);
    const int r = k - g.N;
    const Combined v = combine_cell(a, src, sf_start(a));
    a.V[k] = v.V; a.Z[k] = v.Z;
    a.Vx[k] = v.Vx; a.Vy[k] = v.Vy;
    a.Zx[k] = v.Zx; a.Zy[k] = v.Zy;
    if (r < 2 * g.nx) {                      // jP / jM ghosts
        a.Nx[k] = v.Nx;
        a.Oy[k] = r < g.nx ? v.Oy : 0.0;
    } else {                                 // iP / iM ghosts
        const int jj = (r - 2 * g.nx) % g.ny;
        a.Nx[k] = 0.0;
        a.Oy[k] = jj == g.ny - 1 ? v.OyOpen : v.Oy;
    }
}

// updateAllFaceDepth (plain, nssolve.c:230) + detectWaterfallLocationSub +
// updateVelocitySub + waterfallCorrectionSub (subgrid.c:1335-1444, 1084-1187)
// for the faces of every Ng index.  Interior flags are recomputed every step
// (the Sub variant resets them before detecting); the ghost flags written by
// the XM/YM tests of boundary cells stay set, as in the reference.
// updateVelocitySub + waterfallCorrectionSub of interior cell c's east and
// north faces (with the waterfall flags), given the cell's combined face
// values v; a pure function of the step's inputs, so a ghost face can
// recompute the face it mirrors
template <class R>
__device__ inline FaceVel face_vel_sub(const StepArgsT<R> &a, int c, const Combined &cv)
{
    const Grid &g = a.g;
    const P &p = a.p;
    const int i = c % g.nx, j = c / g.nx;
    const int e = nbE(g, c, i, j), n = nbN(g, c, i, j);
    const double *hs = sf_new(a);
    const double hc = hs[c], he = hs[e], hn = hs[n];
    const double bXc = a.bottomXP[c], bYc = a.bottomYP[c];
    FaceVel r{0, 0, 0, 0, 0};
    double &fX = r.fX, &fY = r.fY;
    if (hc < bXc - 0.0 && he - bXc >= p.wtfh) fX = -1;
    if (e < g.N && he < bXc - 0.0 && hc - bXc >= p.wtfh) fX = 1;
    if (hc < bYc - 0.0 && hn - bYc >= p.wtfh) fY = -1;
    if (n < g.N && hn < bYc - 0.0 && hc - bYc >= p.wtfh) fY = 1;
    const double Nxc = cv.Nx, Oyc = cv.Oy, Vxc = cv.Vx, Vyc = cv.Vy;
    const double effHx = Vxc > 0 ? Nxc / Vxc : 0;
    const double effHy = Vyc > 0 ? Oyc / Vyc : 0;
    const double divider = 1.0;
    double uc = a.EnXP[c] - (p.g * p.dt * effHx) * (he - hc) / divider;
    double vc = a.EnYP[c] - (p.g * p.dt * effHy) * (hn - hc) / divider;
    if (Nxc == 0) uc = 0;
    if (Oyc == 0) vc = 0;
    // no outflow from a dry cell: the face's own cell stops positive flow,
    // the cell beyond stops negative flow
    // (the neighbours' depth and rmv flags loaded up front, not behind the
    // velocity tests; depth and rmv hold interior cells only)
    const double *dn = depth_nxt(a);
    const bool eIn = e < g.N, nIn = n < g.N;
    const double dC = dn[c], dE = eIn ? dn[e] : 1.0, dN = nIn ? dn[n] : 1.0;
    const uint8_t rC = a.rmv[c], rE = eIn ? a.rmv[e] : 0, rN = nIn ? a.rmv[n] : 0;
    const bool dryC = dC == 0;
    if (dryC && uc > 0) uc = 0;
    if (eIn && dE == 0 && uc < 0) uc = 0;
    if (dryC && vc > 0) vc = 0;
    if (nIn && dN == 0 && vc < 0) vc = 0;
    const bool rmvC = rC == 1;
    if (rmvC || (eIn && rE == 1)) uc = 0;
    if (rmvC || (nIn && rN == 1)) vc = 0;
    const double CFLmax = 0.24 / p.subR, Cw = 0.7;
    if (fX != 0) {
        const double vel0 = uc;
        uc = fX * Cw * sqrt(2 * p.g * Nxc / p.dy);
        if (fabs(uc) > CFLmax * p.dx / p.dt) uc = fX * CFLmax * p.dx / p.dt;
        r.loss -= (vel0 - uc) * p.dt * Nxc;
    }
    if (fY != 0) {
        const double vel0 = vc;
        vc = fY * Cw * sqrt(2 * p.g * Oyc / p.dx);
        if (fabs(vc) > CFLmax * p.dy / p.dt) vc = fY * CFLmax * p.dy / p.dt;
        r.loss -= (vel0 - vc) * p.dt * Oyc;
    }
    r.u = uc;
    r.v = vc;
    return r;
}

// combineSubArea of one Ng index (an interior cell, or a ghost recomputing
// the cell it mirrors; the four corner ghosts are not written), then its
// face depths, velocities, waterfall flags and flow rates
// (computeFaceFlowRateSub, subgrid.c:1190-1212).  The face velocities of a
// cell need only that cell's combined values, so one pass does both; a
// ghost face takes the velocity of the face it mirrors (enforceVelocityBC),
// recomputed here from its source cell's combined values.
template <class R>
__device__ inline void combine_faces_cell(const StepArgsT<R> &a, int k, double &loss)
{
    const Grid &g = a.g;
    const P &p = a.p;
    const double dX = clipped(face_depth_x(a, k));
    const double dY = clipped(face_depth_y(a, k));
    double u, v, NxK, OyK;
    if (k < g.N) {
        const Combined cv = combine_cell(a, k, sf_new(a));
        a.Nx[k] = cv.Nx; a.Oy[k] = cv.Oy; a.Vx[k] = cv.Vx; a.Vy[k] = cv.Vy;
        a.Zx[k] = cv.Zx; a.Zy[k] = cv.Zy; a.V[k] = cv.V; a.Z[k] = cv.Z;
        const FaceVel f = face_vel_sub(a, k, cv);
        a.wtfXP[k] = f.fX;
        a.wtfYP[k] = f.fY;
        u = f.u;
        v = f.v;
        NxK = cv.Nx;
        OyK = cv.Oy;
        loss += f.loss;
    } else {
        const int r = k - g.N;
        const int sv = ghost_source(g, k, false), su = ghost_source(g, k, true);
        const Combined cvv = combine_cell(a, sv, sf_new(a));
        if (k < g.Ng - 4) {
            // combineSubArea's ghost copies (subgrid.c:480-506); the iM / iP
            // ghosts of the top row take the corner cell's Oy before the
            // top-row closure
            a.V[k] = cvv.V; a.Z[k] = cvv.Z;
            a.Vx[k] = cvv.Vx; a.Vy[k] = cvv.Vy;
            a.Zx[k] = cvv.Zx; a.Zy[k] = cvv.Zy;
            if (r < 2 * g.nx) {                      // jP / jM ghosts
                NxK = cvv.Nx;
                OyK = r < g.nx ? cvv.Oy : 0.0;
            } else {                                 // iP / iM ghosts
                const int jj = (r - 2 * g.nx) % g.ny;
                NxK = 0.0;
                OyK = jj == g.ny - 1 ? cvv.OyOpen : cvv.Oy;
            }
            a.Nx[k] = NxK;
            a.Oy[k] = OyK;
        } else {
            NxK = a.Nx[k];
            OyK = a.Oy[k];
        }
        if (r >= g.nx && r < 2 * g.nx) {             // jM ghost: YM test of cell (i, 0)
            const int i = r - g.nx;
            const double bot = a.bottomYP[k];
            if (sf_new(a)[i] < bot - 0.0 && sf_new(a)[k] - bot >= p.wtfh) a.wtfYP[k] = 1;
        } else if (r >= 2 * g.nx + g.ny && r < 2 * g.nx + 2 * g.ny) {   // iM ghost
            const int jj = r - 2 * g.nx - g.ny;
            const double bot = a.bottomXP[k];
            if (sf_new(a)[jj * g.nx] < bot - 0.0 && sf_new(a)[k] - bot >= p.wtfh) a.wtfXP[k] = 1;
        }
        u = face_vel_sub(a, su, su == sv ? cvv : combine_cell(a, su, sf_new(a))).u;
        v = face_vel_sub(a, sv, cvv).v;
    }
    a.depthXP[k] = dX;
    a.depthYP[k] = dY;
    a.uuXP[k] = u;
    a.vvYP[k] = v;
    a.Fuu[k] = u * NxK;
    a.Fvv[k] = v * OyK;
}

// combineSubArea + updateAllFaceDepth (plain, nssolve.c:230) +
// detectWaterfallLocationSub + updateVelocitySub + waterfallCorrectionSub
// (subgrid.c:254-526, 1335-1444, 1084-1187) for the faces of every Ng index.
// Interior flags are recomputed every step (the Sub variant resets them
// before detecting); the ghost flags written by the XM/YM tests of boundary
// cells stay set, as in the reference.
template <class R>
__global__ void __launch_bounds__(256) k_combine_faces(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.nb1) return;
    __shared__ double lds[16];
    if (tail_off(a)) return;
    // blocks from the last index down: the ghost faces, which recompute the
    // faces they mirror, in the first blocks, not in the last ones
    const int blk = a.nb1 - 1 - (int)blockIdx.x;
    const int k = blk * blockDim.x + threadIdx.x;
    double loss[1] = {0.0};
    if (k < a.g.Ng) combine_faces_cell(a, k, loss[0]);
    block_sum<1>(loss, lds);
    if (threadIdx.x == 0) a.vpart[3 * a.maxBlocks + blk] += loss[0];
}

// velocityInterp + volumeByFluxSub + updateCDSub (useSubDrag 0: h = V/Z;
// 2: Volp; 3, 4: Casas)
// + thinLayerDrag + CFL numbers (nsfunctions.c:337-347, subgrid.c:1227-1263,
// 1520-1535, nsfunctions.c:691-727, 1192-1232)
template <class R>
__global__ void __launch_bounds__(256) k_finish_sub(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.nb1) return;
    __shared__ double lds[16 * 5];
    if (tail_off(a)) return;
    const Grid g = a.g;
    // blocks start at the tidal boundary (row_block), whose faces take
    // extra work here
    const int blk = a.rowsDown ? a.nb1 - 1 - (int)blockIdx.x : (int)blockIdx.x;
    const int c = blk * blockDim.x + threadIdx.x;
    const P &p = a.p;
    double cflx = 0, cfly = 0;
    VolBudget vb{{0.0, 0.0}, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (c < g.N) {
        const int tt = cur_tt(a);
        const int nx = g.nx, i = c % nx, j = c / nx;
        const auto u = a.uuXP;
        const double um1 = c - 1 >= 0 ? u[c - 1] : 0.0;
        const double um1n = c - 1 + nx >= 0 ? u[c - 1 + nx] : 0.0;
        const InterpV v = interp_v<R, true>(a, c, i, j);
        a.uuYP[c] = 0.25 * (u[c] + um1 + u[c + nx] + um1n);
        a.vvXP[c] = 0.25 * (v.c + v.c1 + v.mn + v.mn1);

        volume_update(a, c, i, j, tt, p.bcType != 3, vb);

        if (p.CDnotN == 0 && p.useSubDrag == 2) {
            // Volp 2013 (subgrid.c:1481-1506): g n^2 Yh, capped, 0 where Yh <= 0
            const double n = a.manning ? a.manning[c] : p.manningN;
            const double coeff = p.g * n * n;
            const double yx = a.YhX[c], yy = a.YhY[c];
            double cx = 0.0, cy = 0.0;
            if (yx > 0) { cx = coeff * yx; if (cx > p.CDmax) cx = p.CDmax; }
            if (yy > 0) { cy = coeff * yy; if (cy > p.CDmax) cy = p.CDmax; }
            a.CDXP[c] = cx;
            a.CDYP[c] = cy;
        } else if (p.CDnotN == 0 && (p.useSubDrag == 3 || p.useSubDrag == 4)) {
            // Casas 2010 (subgrid.c:1508-1519): g Cn^2 / h^(1/3), h = V/(dx dy)
            if (a.V[c] > 0) {
                const double hh = a.V[c] / (p.dx * p.dy);
                const double cn = a.Cn[c];
                const double cd = p.g * cn * cn / pow(hh, 1.0 / 3.0);
                a.CDXP[c] = cd;
                a.CDYP[c] = cd;
            }
        } else if (p.CDnotN == 0 && a.V[c] > 0) {
            const double n = a.manning ? a.manning[c] : p.manningN;
            const double coeff = p.g * n * n;
            const double hh = a.V[c] / a.Z[c];
            const double cd = coeff / pow(hh, 1.0 / 3.0);
            a.CDXP[c] = cd;
            a.CDYP[c] = cd;
        }
        const double d = depth_nxt(a)[c];
        if (p.useThinLayer == 1 && p.CDnotN == 0 && d > 0) {
            const double n = a.manning ? a.manning[c] : p.manningN;
            const double coeff = p.g * n * n;
            if (a.depthXP[c] <= p.hD) a.CDXP[c] = coeff / pow(d, 2.0 / 3.0);
            if (a.depthYP[c] <= p.hD) a.CDYP[c] = coeff / pow(d, 2.0 / 3.0);
        }
        cflx = fabs(u[c]) * p.dt / p.dx;
        cfly = fabs(v.c) * p.dt / p.dy;
        a.CFLx[c] = cflx;
        a.CFLy[c] = cfly;
    }
    if (p.useThinLayer == 1 && p.CDnotN == 1 && c < g.Ng) {
        if (a.depthXP[c] <= p.hD) a.CDXP[c] = p.CDx + (p.CDmax - p.CDx) * (p.hD - a.depthXP[c]) / p.hD;
        if (a.depthYP[c] <= p.hD) a.CDYP[c] = p.CDy + (p.CDmax - p.CDy) * (p.hD - a.depthYP[c]) / p.hD;
    }
    const double mx = block_max(cflx, lds);
    const double my = block_max(cfly, lds);
    if (threadIdx.x == 0) {
        a.cflpart[blk] = mx;
        a.cflpart[a.maxBlocks + blk] = my;
        cfl_stop_test(a, mx, my);
    }
    if (p.budget) store_vol_budget(a, vb, lds, blk);
}

// matrixSourceTermSub + matrixCoeffSub (useSubgrid = 1, staggeredV = 1) +
// adjustMatrixBoundary + adjustTidalBoundary (subgrid.c:930-1081,
// nsfunctions.c:846-924) and the solver start.  A face coefficient is
// g dt^2 N^2 / V / (1 + Drag) (0 where V = 0); the closed wall faces have
// N = 0, so the wall adjustment subtracts zeros.  GnCt equal to 0 or dx*dy
// becomes dx*dy; a dry cell with no velocity on its four faces gets the
// identity-like row dx*dy * x = dx*dy * eta (flagged by a negative D).
template <class R>
__device__ inline Row assemble_row_sub(const StepArgsT<R> &a, int c, int i, int j, int tt, const RowEn &q)
{
    const Grid &g = a.g;
    const P &p = a.p;
    const double dt = p.dt, dxdy = p.dx * p.dy;
    const double cx = p.g * (dt * dt), cy = p.g * (dt * dt);
    const double divider = 1.0;
    const int w = nbW(g, c, i, j), s = nbS(g, c, i, j);
    const double Nxc = a.Nx[c], Nxw = a.Nx[w], Oyc = a.Oy[c], Oys = a.Oy[s];
    const double Vxc = a.Vx[c], Vxw = a.Vx[w], Vyc = a.Vy[c], Vys = a.Vy[s];
    double GXP = 0, GXM = 0, GYP = 0, GYM = 0;
    if (Vxc > 0) GXP = cx * Nxc * Nxc / Vxc / divider / (1 + q.DXc);
    if (Vxw > 0) GXM = cx * Nxw * Nxw / Vxw / divider / (1 + q.DXw);
    if (Vyc > 0) GYP = cy * Oyc * Oyc / Vyc / divider / (1 + q.DYc);
    if (Vys > 0) GYM = cy * Oys * Oys / Vys / divider / (1 + q.DYs);
    double D = a.Z[c] + GXP + GYP + GXM + GYM;
    const double hc = sf_start(a)[c];
    double z = hc * a.Z[c] - dt * (q.EnXc * Nxc - q.EnXw * Nxw) - dt * (q.EnYc * Oyc - q.EnYs * Oys);
    const uint8_t fl = a.flags[c];
    const int ninflow = p.bcType != 3 ? (fl >> 2) : 0;
    for (int k = 0; k < ninflow; k++) z = z + dt * (a.inflow[tt] / p.inflowLocLength);
    if (z == 0 || z == hc * p.dx * p.dy) z = hc * p.dx * p.dy;
    if (D == 0.0 || D == dxdy) D = dxdy;
    if (depth_cur(a)[c] == 0 && a.uuXP[c] == 0 && a.uuXP[w] == 0 && a.vvYP[c] == 0 && a.vvYP[s] == 0) {
        D = -dxdy;
        z = hc * p.dx * p.dy;
    }
    if ((fl & 1) && p.bcType != 2) { z = a.tideP[tt] + p.offset; D = -1.0; }
    if ((fl & 2) && p.bcType != 1) { z = a.tideM[tt] + p.offset; D = -1.0; }
    return Row{GXP, GYP, D, z};
}

// the assembly over the cells, grid-stride over nbRes blocks, reading the
// predictor's En and Drag back (plain grid and subgrid)
template <class R, bool SUB>
__device__ inline void assemble_body(const StepArgsT<R> &a, double *lds)
{
    double part[3] = {0.0, 0.0, 0.0};   // r.r, r.z, b.b
    if (!a.ctl->halted) {
        const Grid g = a.g;
        const int tt = cur_tt(a);
        for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g.N; c += a.nbAsm * blockDim.x) {
            const int i = c % g.nx, j = c / g.nx;
            const int w = nbW(g, c, i, j), s = nbS(g, c, i, j);
            RowEn q;
            q.EnXc = a.EnXP[c];
            q.EnXw = a.EnXP[w];
            q.EnYc = a.EnYP[c];
            q.EnYs = a.EnYP[s];
            q.DXc = a.DragXP[c];
            q.DXw = a.DragXP[w];
            q.DYc = a.DragYP[c];
            q.DYs = a.DragYP[s];
            const Row r = SUB ? assemble_row_sub(a, c, i, j, tt, q) : assemble_row(a, c, i, j, tt, q);
            store_row(a, c, r, part);
        }
    }
    if (a.nbAsm != a.nbRes) return;   // (no partials wanted)
    block_sum<3>(part, lds);
    store_partials<3>(a.partA, a.nbRes, {part[0], part[1], part[2]});
}

template <class R>
__global__ void __launch_bounds__(256) k_assemble(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.nbAsm) return;
    __shared__ double lds[16 * 3];
    assemble_body<R, false>(a, lds);
}

template <class R>
__global__ void __launch_bounds__(256) k_assemble_sub(const StepArgsT<R> *__restrict__ A)
{
    const StepArgsT<R> &a = A[blockIdx.z];
    if (blockIdx.x >= a.nbAsm) return;
    __shared__ double lds[16 * 3];
    assemble_body<R, true>(a, lds);
}

// both storage precisions of every step kernel
template __global__ void k_assemble<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_assemble<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_assemble_sub<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_assemble_sub<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_combine<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_combine<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_combine_faces<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_combine_faces<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_faces<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_faces<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_finish<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_finish<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_finish_sub<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_finish_sub<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_limiter<double, false>(const StepArgsT<double> *__restrict__);
template __global__ void k_limiter<double, true>(const StepArgsT<double> *__restrict__);
template __global__ void k_limiter<float, false>(const StepArgsT<float> *__restrict__);
template __global__ void k_limiter<float, true>(const StepArgsT<float> *__restrict__);
template __global__ void k_limiter_sub<double, false, false>(const StepArgsT<double> *__restrict__);
template __global__ void k_limiter_sub<double, false, true>(const StepArgsT<double> *__restrict__);
template __global__ void k_limiter_sub<double, true, false>(const StepArgsT<double> *__restrict__);
template __global__ void k_limiter_sub<double, true, true>(const StepArgsT<double> *__restrict__);
template __global__ void k_limiter_sub<float, false, false>(const StepArgsT<float> *__restrict__);
template __global__ void k_limiter_sub<float, false, true>(const StepArgsT<float> *__restrict__);
template __global__ void k_limiter_sub<float, true, false>(const StepArgsT<float> *__restrict__);
template __global__ void k_limiter_sub<float, true, true>(const StepArgsT<float> *__restrict__);
template __global__ void k_predictor<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_predictor<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_predictor_sub<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_predictor_sub<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_scalar<double, false>(const StepArgsT<double> *__restrict__);
template __global__ void k_scalar<double, true>(const StepArgsT<double> *__restrict__);
template __global__ void k_scalar<float, false>(const StepArgsT<float> *__restrict__);
template __global__ void k_scalar<float, true>(const StepArgsT<float> *__restrict__);
template __global__ void k_scalar_sub<double, false>(const StepArgsT<double> *__restrict__);
template __global__ void k_scalar_sub<double, true>(const StepArgsT<double> *__restrict__);
template __global__ void k_scalar_sub<float, false>(const StepArgsT<float> *__restrict__);
template __global__ void k_scalar_sub<float, true>(const StepArgsT<float> *__restrict__);
template __global__ void k_step_end<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_step_end<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_subarea_init<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_subarea_init<float>(const StepArgsT<float> *__restrict__);
template __global__ void k_surface<double>(const StepArgsT<double> *__restrict__);
template __global__ void k_surface<float>(const StepArgsT<float> *__restrict__);

}  // namespace frehd
