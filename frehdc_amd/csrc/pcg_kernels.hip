// pcg_kernels.hip -- the Jacobi-preconditioned conjugate gradient of the
// free-surface solve (CGIter, LASpack itersolv.c:355-456, with
// JacobiPrecond precond.c:26-46: z = r * (1/|diag|)), two kernels per
// iteration:
//
//   k_pcg_dir(it)  stop test on the residual left by iteration it-1
//                  (RTCResult rtc.c:48-72, checked before every iteration,
//                  itersolv.c:407-409); x += alpha_{it-1} p_{it-1}
//                  (itersolv.c:418, deferred from iteration it-1 to here so
//                  that p_{it-1} is read once); p_it = z + beta p_{it-1};
//                  q = A p_it (Mul_QV row-wise, operats.c:870-888); block
//                  partials of p.q.
//   k_pcg_res(it)  alpha = rho / p.q; r -= alpha q (itersolv.c:419-423,
//                  LASpack's IsOne(alpha) shortcut kept); block partials
//                  of r.r and r.z for the next stop test and beta.
//
// Every block of a consumer kernel re-reduces all block partials of the
// producer in the same fixed order, so the scalars are bitwise identical in
// every block and from run to run, without atomics.
//
// k_pcg_dir walks a tile of 256 columns x H rows: each wave owns 64
// adjacent columns and slides down the rows, keeping p of rows j-1, j, j+1
// in registers, so every array is read once per iteration (east/west
// neighbours come from the adjacent lanes by shuffle; only lanes 0 and 63
// read one halo column).  Loads run one row ahead of the arithmetic.
#include "frehd_dev.h"
#include "step_kernels.h"

namespace frehd {

template <typename T>
__device__ inline T jacobi_z(T r, T D)
{
    return r * (T(1) / fabs(D));
}

template <typename T>
__device__ inline T direction(T z, T po, T beta, int first, int betaOne)
{
    if (first) return z;
    return betaOne ? z + po : z + beta * po;
}

// raw values of one cell needed to form its new direction: the residual
// (Jacobi: z = r / |D| is formed from it) or, with SSOR, z itself
template <typename T>
struct Raw {
    T r, D, po;
};

template <typename T>
__device__ inline Raw<T> load_raw(const PcgArgs<T> &a, const T *pOld, int c, int first)
{
    Raw<T> v;
    v.r = a.ssor ? a.z[c] : a.r[c];
    v.D = a.D[c];
    v.po = first ? T(0) : pOld[c];
    return v;
}

template <typename T>
__device__ inline T precond_z(const PcgArgs<T> &a, const Raw<T> &v)
{
    return a.ssor ? v.r : jacobi_z(v.r, v.D);
}

template <typename T>
__global__ void __launch_bounds__(256) k_pcg_dir(const PcgArgs<T> *__restrict__ A, int it)
{
    const PcgArgs<T> &a = A[blockIdx.z];
    if ((int)blockIdx.x >= a.nbDir) return;
    __shared__ double lds[16 * 3];
    const T *pOld = a.pBuf[it & 1];
    T *pNew = a.pBuf[(it + 1) & 1];
    PcgScalars *sc = a.sc;
    if (*a.halted) return;
    if (a.outerDone && *a.outerDone >= 0) return;
    const int doneAt = sc->doneAt;
    if (doneAt >= 0 && doneAt < it) return;

    double rs[2];
    reduce_partials<2>(a.partA, it == 0 ? a.nbStart : a.nbRes, rs, lds);
    // r.z: of the Jacobi z from the partials, or k_ssor's sum
    const double rr = rs[0], rz = a.ssor ? sc->rhoHist[it & 1] : rs[1];
    const double rNorm = sqrt(rr);
    const double bNorm = it == 0 ? rNorm : sc->bNorm;
    const double eps = a.epsDyn ? *a.epsDyn : a.eps;
    const bool stop = rNorm < eps * bNorm || (LAS_IS_ZERO(bNorm) && LAS_IS_ONE(1.0 + rNorm)) ||
                      it >= a.maxIter;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (it == 0) { sc->bb = rr; sc->bNorm = bNorm; }
        sc->rr = rr;
        sc->rNorm = rNorm;
        if (!a.ssor) sc->rhoHist[it & 1] = rz;
        sc->iter = it;
        if (stop) sc->doneAt = it;
    }
    const int first = it == 0;
    const double alphaPrev = first ? 0.0 : sc->alpha;
    const int alphaOne = LAS_IS_ONE(alphaPrev);
    const T al = (T)alphaPrev;

    int lb = blockIdx.x;
    if (a.remap) lb = (lb & 7) * (a.nbDir >> 3) + (lb >> 3);
    const int bx = lb % a.gx, by = lb / a.gx;
    double part[1] = {0.0};
    if (by < a.gy) {
        const int nx = a.nx, ny = a.ny;
        const int lane = threadIdx.x & 63;
        const int i0 = bx * 256 + (threadIdx.x & ~63);
        const int i = i0 + lane;
        const bool valid = i < nx;
        const int j0 = by * a.H, j1 = min(j0 + a.H, ny);
        if (stop) {
            // converged: only the deferred x update of iteration it-1 (x = 0
            // when the start residual already passes)
            if (valid)
                for (int j = j0; j < j1; j++) {
                    const int c = j * nx + i;
                    if (first) {
                        a.x[c] = T(0);
                    } else {
                        const T po = pOld[c];
                        a.x[c] = alphaOne ? a.x[c] + po : a.x[c] + al * po;
                    }
                }
        } else {
            const double betaD = first ? 0.0 : rz / sc->rhoHist[(it - 1) & 1];
            const int betaOne = LAS_IS_ONE(betaD);
            const T beta = (T)betaD;
            const int ih = lane == 0 ? i0 - 1 : (lane == 63 ? i0 + 64 : -1);
            const bool halo = ih >= 0 && ih < nx;

            // row j0-1 (south neighbour of the first row)
            T pS = 0, cyS = 0;
            if (valid && j0 > 0) {
                const int c = (j0 - 1) * nx + i;
                const Raw<T> v = load_raw(a, pOld, c, first);
                pS = direction(precond_z(a, v), v.po, beta, first, betaOne);
                cyS = a.cY[c];
            }
            // row j0
            Raw<T> cur = {0, 1, 0}, nxt = {0, 1, 0}, hcur = {0, 1, 0};
            T cxc = 0, cyc = 0, xc = 0;
            if (valid) {
                const int c = j0 * nx + i;
                cur = load_raw(a, pOld, c, first);
                cxc = a.cX[c];
                cyc = j0 < ny - 1 ? a.cY[c] : T(0);
                if (!first) xc = a.x[c];
                if (j0 + 1 < ny) nxt = load_raw(a, pOld, c + nx, first);
            }
            if (halo) hcur = load_raw(a, pOld, j0 * nx + ih, first);
            T pc = direction(precond_z(a, cur), cur.po, beta, first, betaOne);

            for (int j = j0; j < j1; j++) {
                const int c = j * nx + i;
                // loads one row ahead: raw of row j+2, the rest of row j+1
                Raw<T> nn = {0, 1, 0}, hn = {0, 1, 0};
                T cxn = 0, cyn = 0, xn = 0;
                if (valid && j + 1 < j1) {
                    if (j + 2 < ny) nn = load_raw(a, pOld, c + 2 * nx, first);
                    cxn = a.cX[c + nx];
                    cyn = j + 1 < ny - 1 ? a.cY[c + nx] : T(0);
                    if (!first) xn = a.x[c + nx];
                }
                if (halo && j + 1 < j1) hn = load_raw(a, pOld, c + nx - i + ih, first);

                const T pn = (valid && j + 1 < ny) ? direction(precond_z(a, nxt), nxt.po, beta, first, betaOne)
                                                   : T(0);
                const T ph = halo ? direction(precond_z(a, hcur), hcur.po, beta, first, betaOne) : T(0);
                T pw = __shfl_up(pc, 1, 64), pe = __shfl_down(pc, 1, 64);
                T cxw = __shfl_up(cxc, 1, 64);
                if (lane == 0) {
                    pw = ph;
                    cxw = (valid && i > 0) ? a.cX[c - 1] : T(0);
                }
                if (lane == 63) pe = ph;
                if (valid) {
                    const T Dc = cur.D;
                    T sum;
                    if (Dc < 0) {
                        sum = -Dc * pc;
                    } else {
                        sum = 0;
                        if (i > 0) sum += (-cxw) * pw;
                        if (j > 0) sum += (-cyS) * pS;
                        sum += Dc * pc;
                        if (j < ny - 1) sum += (-cyc) * pn;
                        if (i < nx - 1) sum += (-cxc) * pe;
                    }
                    pNew[c] = pc;
                    a.q[c] = sum;
                    part[0] += (double)pc * (double)sum;
                    a.x[c] = first ? T(0) : (alphaOne ? xc + cur.po : xc + al * cur.po);
                }
                pS = pc;
                cyS = cyc;
                pc = pn;
                cur = nxt;
                nxt = nn;
                hcur = hn;
                cxc = cxn;
                cyc = cyn;
                xc = xn;
            }
        }
    }
    block_sum<1>(part, lds);
    if (!stop) store_partials<1>(a.partB, a.nbDir, {part[0]});
}

// ---------------------------------------------------------------------------
// SSORPrecond with omega = 1 (precond.c:48-73): z = (D+U)^-1 D (D+L)^-1 r,
// the rows in the solver's transposed order k = i*ny + j (cell j*nx + i).
// The forward solve with D+L (MulInv_QV, operats.c:1136-1160) takes row k
// after rows k-ny (west) and k-1 (south); the backward solve with D+U
// (operats.c:1081-1105) after k+ny (east) and k+1 (north).  Every cell of
// an anti-diagonal i + j = d depends only on the diagonal before it, so the
// sweeps run diagonal by diagonal with a barrier in between, each row
// evaluated with the reference's own expression and order (west then south;
// east then north; the diagonal product of Mul_QV in between), so z is the
// reference's bit for bit.  Identity rows (D < 0: tidal Dirichlet rows,
// isolated dry subgrid cells) have zero off-diagonals, as the reference's
// adjusted rows do.  One workgroup per member: nx + ny - 1 dependent stages
// per sweep.  The sweep's vector lives in LDS when it fits, else in z.
template <class V>
__device__ inline void ssor_sweeps(const PcgArgs<double> &a, V t)
{
    const int nx = a.nx, ny = a.ny, nd = nx + ny - 1;
    for (int d = 0; d < nd; d++) {
        const int ilo = max(0, d - (ny - 1)), ihi = min(d, nx - 1);
        for (int i = ilo + (int)threadIdx.x; i <= ihi; i += blockDim.x) {
            const int j = d - i, c = j * nx + i;
            const double D = a.D[c];
            const bool id = D < 0;
            double sum = 0.0;
            if (i > 0) sum -= (-(id ? 0.0 : a.cX[c - 1])) * t[c - 1];
            if (j > 0) sum -= (-(id ? 0.0 : a.cY[c - nx])) * t[c - nx];
            sum += a.r[c];
            t[c] = sum * (1.0 / fabs(D));
        }
        __syncthreads();
    }
    for (int d = nd - 1; d >= 0; d--) {
        const int ilo = max(0, d - (ny - 1)), ihi = min(d, nx - 1);
        for (int i = ilo + (int)threadIdx.x; i <= ihi; i += blockDim.x) {
            const int j = d - i, c = j * nx + i;
            const double D = a.D[c];
            const bool id = D < 0;
            double sum = 0.0;
            if (i < nx - 1) sum -= (-(id ? 0.0 : a.cX[c])) * t[c + 1];
            if (j < ny - 1) sum -= (-(id ? 0.0 : a.cY[c])) * t[c + nx];
            double m = 0.0;
            m += fabs(D) * t[c];
            sum += m;
            t[c] = sum * (1.0 / fabs(D));
        }
        __syncthreads();
    }
}

template <bool LDS>
__global__ void __launch_bounds__(1024) k_ssor(const PcgArgs<double> *__restrict__ A, int it)
{
    const PcgArgs<double> &a = A[blockIdx.z];
    extern __shared__ double zl[];
    __shared__ double lds[16];
    if (!a.ssor || *a.halted) return;
    PcgScalars *sc = a.sc;
    const int doneAt = sc->doneAt;
    if (doneAt >= 0 && doneAt < it) return;
    if (LDS) ssor_sweeps(a, zl);
    else ssor_sweeps(a, a.z);
    double part[1] = {0.0};
    for (int c = threadIdx.x; c < a.N; c += blockDim.x) {
        const double z = LDS ? zl[c] : (double)a.z[c];
        if (LDS) a.z[c] = z;
        part[0] += a.r[c] * z;
    }
    block_sum<1>(part, lds);
    if (threadIdx.x == 0) sc->rhoHist[it & 1] = part[0];
}

template <typename T>
struct alignas(16) Pack16 {
    T v[16 / sizeof(T)];
};

template <typename T>
__global__ void __launch_bounds__(256) k_pcg_res(const PcgArgs<T> *__restrict__ A, int it)
{
    const PcgArgs<T> &a = A[blockIdx.z];
    if ((int)blockIdx.x >= a.nbRes) return;
    __shared__ double lds[16 * 3];
    PcgScalars *sc = a.sc;
    if (*a.halted) return;
    if (a.outerDone && *a.outerDone >= 0) return;
    const int doneAt = sc->doneAt;
    if (doneAt >= 0 && doneAt <= it) return;
    double pqv[1];
    reduce_partials<1>(a.partB, a.nbDir, pqv, lds);
    const double alpha = sc->rhoHist[it & 1] / pqv[0];
    const int alphaOne = LAS_IS_ONE(alpha);
    const T al = (T)alpha;
    constexpr int V = 16 / sizeof(T);
    double part[2] = {0.0, 0.0};
    auto cell = [&](T &rk, T qk, T Dk) {
        rk = alphaOne ? rk - qk : rk - al * qk;
        part[0] += (double)rk * (double)rk;
        part[1] += (double)rk * (double)(rk * (T(1) / fabs(Dk)));
    };
    const int nv = a.N / V;
    const Pack16<T> *q2 = reinterpret_cast<const Pack16<T> *>(a.q);
    const Pack16<T> *D2 = reinterpret_cast<const Pack16<T> *>(a.D);
    Pack16<T> *r2 = reinterpret_cast<Pack16<T> *>(a.r);
    for (int h = blockIdx.x * blockDim.x + threadIdx.x; h < nv; h += a.nbRes * blockDim.x) {
        const Pack16<T> qv = q2[h], dv = D2[h];
        Pack16<T> rv = r2[h];
#pragma unroll
        for (int u = 0; u < V; u++) cell(rv.v[u], qv.v[u], dv.v[u]);
        r2[h] = rv;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int k = nv * V; k < a.N; k++) {
            T rk = a.r[k];
            cell(rk, a.q[k], a.D[k]);
            a.r[k] = rk;
        }
    block_sum<2>(part, lds);
    store_partials<2>(a.partA, a.nbRes, {part[0], part[1]});
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->pq = pqv[0];
        sc->alpha = alpha;
        sc->iter = it + 1;
    }
}

// ---------------------------------------------------------------------------
// Mixed precision: FP64 iterative refinement around the FP32 solve.
//
// k_refine(s): x_s = x_{s-1} + S d (s > 0; x_0 is the start the assembly
// wrote; S = |D|^-1/2) and the true residual r = b - A x_s with the FP64
// operator, in the row order of the operator (W, S, C, N, E); the same
// sliding-window walk as k_pcg_dir, over x_s recomputed at the neighbours.
// The inner solve runs on the decoupled operator (rows ignore identity-row
// neighbours), so its right-hand side is lifted: a row next to an identity
// row t gets + G * d_t, with d_t = r_t / (-D_t) the identity row's exact
// correction; the decoupled system then has the coupled system's solution.
// It is also Jacobi-scaled, S A S y = S (r + lift), d = S y with S about
// |D|^-1/2: unit diagonal up to the FP32 rounding of S, so the inner CG
// needs no preconditioner and reads no diagonal.  Writes
// x_s, r32 = (float) S (r + lift), at s = 0 the scaled FP32 coefficients,
// and the partials b.b (s = 0), r.r (the FP64 true residual) and r32.r32
// (the inner solve's first stop test).
// The scaling S of the inner system, per cell from its diagonal: about
// |D|^-1/2 (one FP32 reciprocal square root).  Any positive diagonal S
// gives the same solution; what matters is that the operator, the
// right-hand side and the correction use the same S of a cell, which this
// one function guarantees.
__device__ inline double diag_scale(double D)
{
    return (double)rsqrtf((float)fabs(D));
}

// refinement sweep s over tile lb (tile column lb % gx, row lb / gx):
// x_s, r32 (and at s = 0 the scaled coefficients) of the tile's cells; the
// tile's sums r32.r32, r32.z32, b.b, r.r added to part
__device__ inline void refine_tile(const RefineArgs &a, int s, int lb, double (&part)[4])
{
    const double *xin = s == 0 ? a.xbuf[0] : a.xbuf[(s + 1) & 1];   // x_0 or x_{s-1}
    double *xout = a.xbuf[s & 1];
    const int bx = lb % a.gx, by = lb / a.gx;
    if (by < a.gy) {
        const int nx = a.nx, ny = a.ny;
        const int lane = threadIdx.x & 63;
        const int i0 = bx * 256 + (threadIdx.x & ~63);
        const int i = i0 + lane;
        const bool valid = i < nx;
        const int j0 = by * a.H, j1 = min(j0 + a.H, ny);
        const int ih = lane == 0 ? i0 - 1 : (lane == 63 ? i0 + 64 : -1);
        const bool halo = ih >= 0 && ih < nx;
        // x_s at a cell from its diagonal: the inner solve returns the
        // scaled correction, d = S y
        auto xs = [&](int c, double Dv) { return s > 0 ? xin[c] + (double)a.d[c] * diag_scale(Dv) : xin[c]; };
        // rows j-1 (S), j (C), j+1 (N): x_s, diagonal and right-hand side
        double xS = 0, DS = 1, bS = 0, cyS = 0;
        if (valid && j0 > 0) {
            const int c = (j0 - 1) * nx + i;
            DS = a.D[c];
            xS = xs(c, DS);
            bS = a.b[c];
            cyS = a.cY[c];
        }
        double xc = 0, Dc = 1, bc = 0, xn = 0, Dn = 1, bn = 0;
        if (valid) {
            const int c = j0 * nx + i;
            Dc = a.D[c];
            xc = xs(c, Dc);
            bc = a.b[c];
            if (j0 + 1 < ny) {
                Dn = a.D[c + nx];
                xn = xs(c + nx, Dn);
                bn = a.b[c + nx];
            }
        }
        for (int j = j0; j < j1; j++) {
            const int c = j * nx + i;
            double xh = 0, Dh = 1, bh = 0;
            if (halo) {
                const int h = j * nx + ih;
                Dh = a.D[h];
                xh = xs(h, Dh);
                bh = a.b[h];
            }
            const double cxc = valid ? a.cX[c] : 0.0;
            double xw = __shfl_up(xc, 1, 64), xe = __shfl_down(xc, 1, 64);
            double Dw = __shfl_up(Dc, 1, 64), De = __shfl_down(Dc, 1, 64);
            double bw = __shfl_up(bc, 1, 64), be = __shfl_down(bc, 1, 64);
            double cxw = __shfl_up(cxc, 1, 64);
            if (lane == 0) {
                xw = xh;
                Dw = Dh;
                bw = bh;
                cxw = (valid && i > 0) ? a.cX[c - 1] : 0.0;
            }
            if (lane == 63) {
                xe = xh;
                De = Dh;
                be = bh;
            }
            // next row (j+2) ahead
            double xnn = 0, Dnn = 1, bnn = 0;
            if (valid && j + 2 < ny && j + 1 < j1) {
                Dnn = a.D[c + 2 * nx];
                xnn = xs(c + 2 * nx, Dnn);
                bnn = a.b[c + 2 * nx];
            }
            double cyc = 0.0;
            if (valid) {
                cyc = j < ny - 1 ? a.cY[c] : 0.0;
                double sum, lift = 0.0;
                if (Dc < 0) {
                    sum = -Dc * xc;
                } else {
                    sum = 0.0;
                    if (i > 0) sum += (-cxw) * xw;
                    if (j > 0) sum += (-cyS) * xS;
                    sum += Dc * xc;
                    if (j < ny - 1) sum += (-cyc) * xn;
                    if (i < nx - 1) sum += (-cxc) * xe;
                    if (i > 0 && Dw < 0) lift += cxw * ((bw + Dw * xw) / -Dw);
                    if (j > 0 && DS < 0) lift += cyS * ((bS + DS * xS) / -DS);
                    if (j < ny - 1 && Dn < 0) lift += cyc * ((bn + Dn * xn) / -Dn);
                    if (i < nx - 1 && De < 0) lift += cxc * ((be + De * xe) / -De);
                }
                const double r = bc - sum;
                // the inner right-hand side, scaled: S (r + lift)
                const double Sc = diag_scale(Dc);
                const float r32 = (float)((r + lift) * Sc);
                if (s > 0) xout[c] = xc;
                a.r32[c] = r32;
                if (s == 0) {
                    // the scaled inner operator: c / sqrt(D_c D_nb), zero on
                    // faces that touch an identity row or the domain edge
                    a.cXs[c] = (i < nx - 1 && Dc > 0 && De > 0) ? (float)(cxc * Sc * diag_scale(De)) : 0.0f;
                    a.cYs[c] = (j < ny - 1 && Dc > 0 && Dn > 0) ? (float)(cyc * Sc * diag_scale(Dn)) : 0.0f;
                }
                part[0] += (double)r32 * (double)r32;
                part[1] += (double)r32 * (double)r32;
                if (s == 0) part[2] += bc * bc;
                part[3] += r * r;
            }
            xS = xc; DS = Dc; bS = bc; cyS = cyc;
            xc = xn; Dc = Dn; bc = bn;
            xn = xnn; Dn = Dnn; bn = bnn;
        }
    }
}

// the block-reduced sums of a refinement tile into partial slot `slot`
// (thread 0)
__device__ inline void refine_store(const RefineArgs &a, int slot, const double (&part)[4])
{
    if (threadIdx.x != 0) return;
    a.partA[slot] = part[0];
    a.partA[a.nb + slot] = part[1];
    a.partC[slot] = part[2];
    a.partC[a.nb + slot] = part[3];
}


// ---------------------------------------------------------------------------
// Inner FP32 solve of the mixed path: CG on the scaled, decoupled operator
// A' = S A S (unit diagonal; off-diagonals -cXs, -cYs, zero on faces that
// touch an identity row or the domain edge), one kernel and one reduction
// per iteration.
//
// k_cg1(it), it >= 1, from r_{it-1}, p_{it-1} and the block partials of
// k_cg1(it-1) (p.q, r.q, q.q, r.r with q = A' p):
//   alpha = r.r / p.q                          (direct sums)
//   rho   = r.r - 2 alpha r.q + alpha^2 q.q     (|r_it|^2 by recurrence)
//   beta  = rho / r.r
//   q_{it-1} = A' p_{it-1}, recomputed here rather than stored
//   r_it = r_{it-1} - alpha q_{it-1}     x_it = x_{it-1} + alpha p_{it-1}
//   p_it = r_it + beta p_{it-1}          q_it = A' p_it -> block partials
// It reads r, p, x, cXs, cYs and writes r, p, x: the 8 accesses per cell of
// one iteration.  r and p alternate between two buffers, because blocks
// read their neighbours' previous values.  k_cg1(0) forms the partials of
// p_0 = r_0 (the scaled residual of k_refine) and zeroes x.  The stop test
// |r| < eps |r_0| runs at the start of a launch on the last directly summed
// r.r (r_0 at it = 0, r_{it-1} from it = 2 on); a stopped solve's x is
// final, and so is the x of a solve the launch pattern cuts short.
//
// Tile walk as in k_pcg_dir (loads two rows ahead), W adjacent columns
// per lane (W = 1, 2, 4: one float, float2 or float4 access per array and
// row; a wave owns 64 W columns, a block 256 W) and slides down the rows,
// ingesting row t (q_{it-1}, r_it, p_it of row t) and finishing row t-1
// (q_it, the stores, the partials).  A' p_{it-1} at a neighbour needs
// p_{it-1} two cells away: lanes 0 and 63 also ingest the column beyond
// the wave (v) from the column after it (o), with the same expressions the
// owner of v uses, so both obtain bitwise the same p_it there.  Every cell's
// values do not depend on W; the partial sums' order does.
namespace {

template <int W>
struct Cg1Row {
    float P2[W], R[W], cX[W], cY[W], X[W];  // own columns: p(t+1), r(t), cXs(t), cYs(t), x(t-1)
    float hP2, hR, hcXv, hcY, hPo, hcXo;    // lanes 0 / 63: column v and its outer neighbour o
};

// clang vector types (not HIP's float2 / float4 classes, whose member-wise
// copies go through generic references and lose the global address space)
typedef float vf2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));
template <int W> struct VecOf;
template <> struct VecOf<1> { typedef float T; };
template <> struct VecOf<2> { typedef vf2 T; };
template <> struct VecOf<4> { typedef vf4 T; };

__device__ inline void unpack(float v, float (&o)[1]) { o[0] = v; }
__device__ inline void unpack(vf2 v, float (&o)[2]) { o[0] = v.x; o[1] = v.y; }
__device__ inline void unpack(vf4 v, float (&o)[4]) { o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }
__device__ inline float pack(const float (&o)[1]) { return o[0]; }
__device__ inline vf2 pack(const float (&o)[2]) { return vf2{o[0], o[1]}; }
__device__ inline vf4 pack(const float (&o)[4]) { return vf4{o[0], o[1], o[2], o[3]}; }

// Buffer loads of the N-float vectors.  The resource covers the array, and
// an offset beyond it reads zeros: a row above or below the grid or a
// column beyond it loads as 0 with no branch.  (A branch around the load
// makes its registers a merge with the zero path, and the copy that merge
// needs waits for the load at once, so the rows loaded ahead would not stay
// in flight.)  Byte offsets are 32-bit: N < 2^29, checked at create.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr uint32_t OFF_NONE = 0x80000000u;

__device__ inline Rsrc rsrc_of(const float *p, int n)
{
    // the pointers are the same in every lane (argument-block fields picked by
    // the iteration number); say so, or the compiler wraps each load in a
    // loop over the distinct resources of the wave
    const uint64_t v = (uint64_t)p;
    const uint64_t u = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    // word 3: 32-bit data format, the raw-buffer setting of the gfx9 family
    return __builtin_amdgcn_make_buffer_rsrc((void *)u, (short)0, __builtin_amdgcn_readfirstlane(n * 4), 0x00020000);
}

__device__ inline float bload1(Rsrc r, uint32_t off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ inline void bload(Rsrc r, uint32_t off, float (&o)[1]) { o[0] = bload1(r, off); }
__device__ inline void bload(Rsrc r, uint32_t off, float (&o)[2])
{
    unpack(__builtin_bit_cast(vf2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0)), o);
}
__device__ inline void bload(Rsrc r, uint32_t off, float (&o)[4])
{
    unpack(__builtin_bit_cast(vf4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0)), o);
}

// byte offset of (row, col), or OFF_NONE outside the grid or when !ok
__device__ inline uint32_t off_of(int row, int col, int nx, int ny, bool ok)
{
    return (ok && row >= 0 && row < ny) ? (uint32_t)(row * nx + col) * 4u : OFF_NONE;
}

// row `row`, columns i .. i+W-1 (i a multiple of W, nx too), or zeros
template <int W>
__device__ inline void load_cols(Rsrc p, int row, int i, int nx, int ny, bool ok, float (&o)[W])
{
    bload(p, off_of(row, i, nx, ny, ok), o);
}

template <int W>
__device__ inline void store_cols(float *p, size_t c, const float (&v)[W])
{
    gst(reinterpret_cast<typename VecOf<W>::T *>(p + c), pack(v));
}

// A' p at one cell, the terms in the owner's order (W, E, S, N)
__device__ inline float apply_scaled(float pc, float cW, float pW, float cE, float pE, float cS, float pS, float cN,
                                     float pN)
{
#pragma clang fp contract(off)
    return pc - cW * pW - cE * pE - cS * pS - cN * pN;
}

// r_it and p_it of a cell from r_{it-1}, q_{it-1}, p_{it-1}
__device__ inline void advance(float R, float q, float P, float al, float be, float &rn, float &pn)
{
#pragma clang fp contract(off)
    rn = R - al * q;
    pn = rn + be * P;
}

}  // namespace

// The start of inner iteration it: the stop test on the partials of the
// launch before (r_0's norm at it = 0; p.q, r.q, q.q, r.r of iteration it-1
// after), alpha and beta.  Every block evaluates it from the same partials
// in the same order, so all obtain bitwise the same result; thread 0 of the
// block with `record` writes the solver's scalars.  bNorm: |r_0| of this
// solve (read at it >= 2).
struct Cg1Dec {
    bool stop;
    double alphaD, betaD, rNorm0;
};

__device__ inline Cg1Dec cg1_decide(const PcgArgs<float> &a, int it, double eps, double bNorm, double *lds,
                                    bool record)
{
    PcgScalars *sc = a.sc;
    Cg1Dec d{false, 0.0, 0.0, 0.0};
    if (it == 0) {
        double v[1];
        reduce_partials<1>(a.partA, a.nbStart, v, lds);
        const double rNorm = sqrt(v[0]);
        d.rNorm0 = rNorm;
        d.stop = (LAS_IS_ZERO(rNorm) && LAS_IS_ONE(1.0 + rNorm)) || a.maxIter <= 0;
        if (record && threadIdx.x == 0) {
            sc->bb = v[0];
            sc->bNorm = rNorm;
            sc->rr = v[0];
            sc->rNorm = rNorm;
            sc->iter = 0;
            if (d.stop) sc->doneAt = 0;
        }
    } else {
        double v[4];
        reduce_partials<4>(a.partF[(it - 1) & 1], a.nbDir, v, lds);
        const double pq = v[0], rq = v[1], qq = v[2], rr = v[3];
        const double rNorm = sqrt(rr);
        const bool stop = (it >= 2 && (rNorm < eps * bNorm || it - 1 >= a.maxIter)) || !(pq > 0);
        const double alphaD = rr / pq;
        double rho = rr - 2.0 * alphaD * rq + alphaD * alphaD * qq;
        if (!(rho > 0)) rho = 0;               // cancellation at convergence: restart the directions
        d.stop = stop;
        d.alphaD = alphaD;
        d.betaD = rho / rr;
        if (record && threadIdx.x == 0) {
            sc->rr = rr;
            sc->rNorm = rNorm;
            sc->iter = stop ? it - 1 : it;
            if (stop) {
                sc->doneAt = it;
            } else {
                sc->pq = pq;
                sc->alpha = alphaD;
            }
        }
    }
    return d;
}

// k_refine(s).  For s > 0, block 0 first makes the stop test the launch
// after the inner solve's last iteration would make (the launch pattern
// ends the inner solve at its last working iteration): an inner solve whose
// test passes there is recorded as converged, as its own launch would have
// recorded it (the solver statistics and the next pattern read this).
__global__ void __launch_bounds__(256) k_refine(const RefineArgs *__restrict__ A, int s,
                                                const PcgArgs<float> *__restrict__ P)
{
    const RefineArgs &a = A[blockIdx.z];
    if ((int)blockIdx.x >= a.nb) return;
    __shared__ double lds[16 * 4];
    Ctl *ctl = a.ctl;
    const int halted = ctl->halted, refDone = ctl->ref.doneAt;
    if (halted || (refDone >= 0 && refDone < s)) return;
    if (s > 0 && blockIdx.x == 0) {
        const PcgArgs<float> &pa = P[blockIdx.z];
        PcgScalars *sc = pa.sc;
        const int doneAt = sc->doneAt, last = sc->iter;
        if (doneAt < 0 && last >= 1) {
            const Cg1Dec d = cg1_decide(pa, last + 1, ctl->ref.epsIn, sc->bNorm, lds, false);
            if (d.stop && threadIdx.x == 0) sc->doneAt = last + 1;
        }
    }
    double part[4] = {0.0, 0.0, 0.0, 0.0};     // r32.r32, r32.z32, b.b, r.r
    refine_tile(a, s, tile_of(blockIdx.x, a.nb, a.remap), part);
    block_sum<4>(part, lds);
    refine_store(a, blockIdx.x, part);
}

// inner iteration it over tile lb (see above): the partials of q_it added to
// part; with stop at it = 0, the tile's correction is zeroed
template <bool First, int W>
__device__ inline void cg1_tile(const PcgArgs<float> &a, int it, float al, float be, bool stop, int lb,
                                double (&part)[4])
{
    constexpr bool first = First;   // it == 0
    const int bx = lb % a.gx, by = lb / a.gx;
    if (by < a.gy && (!stop || first)) {
        const int nx = a.nx, ny = a.ny;
        const int lane = threadIdx.x & 63;
        const int i0 = bx * 256 * W + (threadIdx.x >> 6) * 64 * W;   // the wave's first column
        const int i = i0 + lane * W;
        const bool valid = i < nx;                                   // nx is a multiple of W
        const int j0 = by * a.H, j1 = min(j0 + a.H, ny);
        // buffers: r and p of iteration it-1 (p_0 = r_0) and of it
        float *const rBuf[2] = {a.r, a.q};
        const float *Rin = first ? a.r : rBuf[(it - 1) & 1];
        const float *Pin = it <= 1 ? a.r : a.pBuf[(it - 1) & 1];
        float *Rout = rBuf[it & 1];
        float *Pout = a.pBuf[it & 1];
        float *X = a.x;
        if (stop) {
            // stopped at it = 0: the correction is zero
            float z[W];
#pragma unroll
            for (int k = 0; k < W; k++) z[k] = 0.0f;
            if (valid)
                for (int j = j0; j < j1; j++) store_cols<W>(X, (size_t)j * nx + i, z);
        } else {
            const Rsrc cXs = rsrc_of(a.cX, a.N), cYs = rsrc_of(a.cY, a.N);
            // halo lanes: v = the column beyond the wave, o = the one after it
            const int v = lane == 0 ? i0 - 1 : (lane == 63 ? i0 + 64 * W : -1);
            const bool halo = v >= 0 && v < nx;
            const int o = lane == 0 ? v - 1 : v + 1;
            const bool oin = halo && o >= 0 && o < nx;
            const Rsrc rIn = rsrc_of(Rin, a.N), pIn = rsrc_of(Pin, a.N), xR = rsrc_of(X, a.N);
            auto at = [&](Rsrc p, int row, int col, bool ok) -> float {
                return bload1(p, off_of(row, col, nx, ny, ok));
            };
            auto load = [&](int t, Cg1Row<W> &L) {
                load_cols<W>(pIn, t + 1, i, nx, ny, valid, L.P2);
                load_cols<W>(rIn, t, i, nx, ny, valid, L.R);
                load_cols<W>(cXs, t, i, nx, ny, valid, L.cX);
                load_cols<W>(cYs, t, i, nx, ny, valid, L.cY);
                load_cols<W>(xR, t - 1, i, nx, ny, valid && !first && t - 1 >= j0, L.X);
                L.hP2 = at(pIn, t + 1, v, halo);
                L.hR = at(rIn, t, v, halo);
                L.hcXv = at(cXs, t, v, halo);
                L.hcY = at(cYs, t, v, halo);
                L.hPo = at(pIn, t, o, oin);
                L.hcXo = at(cXs, t, o, oin && lane == 0);
            };
            // window at the start: rows j0-2 and j0-1 of p, the south face of row j0-1
            float P0[W], P1[W], cYm[W], cYmm[W], cXc[W], pnm[W], pnc[W], rnc[W];
            load_cols<W>(pIn, j0 - 2, i, nx, ny, valid, P0);
            load_cols<W>(pIn, j0 - 1, i, nx, ny, valid, P1);
            load_cols<W>(cYs, j0 - 2, i, nx, ny, valid, cYm);
#pragma unroll
            for (int k = 0; k < W; k++) cYmm[k] = cXc[k] = pnm[k] = pnc[k] = rnc[k] = 0.0f;
            float hP0 = at(pIn, j0 - 2, v, halo), hP1 = at(pIn, j0 - 1, v, halo);
            float hcYm = at(cYs, j0 - 2, v, halo), hpn = 0.0f, hcVIprev = 0.0f;
            // loads run two rows ahead of the arithmetic, into three row
            // buffers that rotate by name (the loop body is instanced three
            // times): a copy of a buffer still being loaded would wait for it
            Cg1Row<W> b0, b1, b2;
            load(j0 - 1, b0);
            load(j0, b1);
            auto row = [&](int t, const Cg1Row<W> &cur, Cg1Row<W> &nxt2) {
                if (t + 2 <= j1) load(t + 2, nxt2);
                // own columns: q_{it-1}, r_it, p_it of row t
                float pWl = __shfl_up(P1[W - 1], 1, 64), pEl = __shfl_down(P1[0], 1, 64);
                float cWl = __shfl_up(cur.cX[W - 1], 1, 64);
                if (lane == 0) { pWl = hP1; cWl = cur.hcXv; }
                if (lane == 63) pEl = hP1;
                float rn[W], pn[W];
#pragma unroll
                for (int k = 0; k < W; k++) {
                    if (first) {
                        rn[k] = cur.R[k];
                        pn[k] = cur.R[k];
                    } else {
                        const float pW = k == 0 ? pWl : P1[k - 1];
                        const float pE = k == W - 1 ? pEl : P1[k + 1];
                        const float cW = k == 0 ? cWl : cur.cX[k - 1];
                        const float q = apply_scaled(P1[k], cW, pW, cur.cX[k], pE, cYm[k], P0[k], cur.cY[k], cur.P2[k]);
                        advance(cur.R[k], q, P1[k], al, be, rn[k], pn[k]);
                    }
                }
                // halo column v of row t, the owner's expressions, evaluated
                // in every lane and kept by lanes 0 and 63 (the other lanes'
                // halo loads read zeros): no divergent branch in the row walk
                float hpnT;
                if (first) {
                    hpnT = cur.hR;
                } else {
                    const float hq0 = apply_scaled(hP1, cur.hcXo, cur.hPo, cur.hcXv, P1[0], hcYm, hP0, cur.hcY, cur.hP2);
                    const float hq63 = apply_scaled(hP1, cur.cX[W - 1], P1[W - 1], cur.hcXv, cur.hPo, hcYm, hP0,
                                                    cur.hcY, cur.hP2);
                    float hr;
                    advance(cur.hR, lane == 0 ? hq0 : hq63, hP1, al, be, hr, hpnT);
                }
                hpnT = halo ? hpnT : 0.0f;
                // finish row j = t-1: q_it = A' p_it, stores, partials
                const int j = t - 1;
                float qWl = __shfl_up(pnc[W - 1], 1, 64), qEl = __shfl_down(pnc[0], 1, 64);
                float cWjl = __shfl_up(cXc[W - 1], 1, 64);
                if (lane == 0) { qWl = hpn; cWjl = hcVIprev; }
                if (lane == 63) qEl = hpn;
                if (valid && j >= j0) {
                    const size_t c = (size_t)j * nx + i;
                    float xs[W];
#pragma unroll
                    for (int k = 0; k < W; k++) {
                        const float qW = k == 0 ? qWl : pnc[k - 1];
                        const float qE = k == W - 1 ? qEl : pnc[k + 1];
                        const float cWj = k == 0 ? cWjl : cXc[k - 1];
                        const float qn = apply_scaled(pnc[k], cWj, qW, cXc[k], qE, cYmm[k], pnm[k], cYm[k], pn[k]);
                        xs[k] = first ? 0.0f : cur.X[k] + al * P0[k];
                        part[0] += (double)pnc[k] * (double)qn;
                        part[1] += (double)rnc[k] * (double)qn;
                        part[2] += (double)qn * (double)qn;
                        part[3] += (double)rnc[k] * (double)rnc[k];
                    }
                    if (!first) {
                        store_cols<W>(Rout, c, rnc);
                        store_cols<W>(Pout, c, pnc);
                    }
                    store_cols<W>(X, c, xs);
                }
                // slide
#pragma unroll
                for (int k = 0; k < W; k++) {
                    P0[k] = P1[k]; P1[k] = cur.P2[k];
                    cYmm[k] = cYm[k]; cYm[k] = cur.cY[k];
                    cXc[k] = cur.cX[k];
                    pnm[k] = pnc[k]; pnc[k] = pn[k]; rnc[k] = rn[k];
                }
                hP0 = hP1; hP1 = cur.hP2;
                hcYm = cur.hcY;
                hpn = hpnT;
                hcVIprev = cur.hcXv;
            };
            for (int t = j0 - 1; t <= j1; t += 3) {
                row(t, b0, b2);
                if (t + 1 > j1) break;
                row(t + 1, b1, b0);
                if (t + 2 > j1) break;
                row(t + 2, b2, b1);
            }
        }
    }
}

template <bool First, int W>
__global__ void __launch_bounds__(256, 3) k_cg1(const PcgArgs<float> *__restrict__ A, int it,
                                             const RefineArgs *__restrict__ R, int testS)
{
    it = First ? 0 : it;
    const PcgArgs<float> &a = A[blockIdx.z];
    if ((int)blockIdx.x >= a.nbDir) return;
    __shared__ double lds[16 * 4];
    PcgScalars *sc = a.sc;
    // the three stop flags in one round trip (most launches of a pattern
    // past convergence end here)
    const int halted = *a.halted;
    int outer = *a.outerDone, doneAt = sc->doneAt;
    double eps = a.epsDyn ? *a.epsDyn : a.eps;
    if (First && testS >= 0) {
        // the outer test of the sweep before this solve, in every thread of
        // every block (testS is uniform; the test reads only the sweep's
        // partials): converged, or this solve's tolerance and a fresh start.
        // Thread 0 of block 0 records it, unless the member has halted or
        // converged before.  A thread that read the control block after that
        // record saw the convergence this test finds, or the unchanged
        // outer flag; the inner flag it replaces.
        const RefTest t = refine_test(R[blockIdx.z], testS, lds, blockIdx.x == 0 && !halted && outer < 0);
        if (t.conv) outer = testS;
        else {
            doneAt = -1;
            eps = t.epsIn;
        }
    }
    if (halted || outer >= 0 || (doneAt >= 0 && doneAt < it)) return;

    const Cg1Dec d = cg1_decide(a, it, eps, sc->bNorm, lds, blockIdx.x == 0);
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    cg1_tile<First, W>(a, it, (float)d.alphaD, (float)d.betaD, d.stop, tile_of(blockIdx.x, a.nbDir, a.remap), part);
    block_sum<4>(part, lds);
    if (!d.stop) store_partials<4>(a.partF[it & 1], a.nbDir, {part[0], part[1], part[2], part[3]});
}

// ---------------------------------------------------------------------------
// FP64 Jacobi PCG in one kernel and one reduction per iteration (the FP64
// path with JacobiPrecond; SSOR keeps the k_pcg_dir / k_pcg_res pair).
// CGIter's recurrences (itersolv.c:396-421, z = r / |D|) regrouped so that
// launch L finishes iteration L-1 and starts iteration L, from r_{L-1},
// p_{L-1}, x_{L-1} and the block partials of launch L-1 (p.q, z.q, q.M^-1 q,
// r.r, r.z of iteration L-1, q = A p):
//   alpha = rho_{L-1} / p.q                 (rho_{L-1} = r.z, summed directly)
//   q_{L-1} = A p_{L-1}, recomputed here rather than stored
//   x_L = x_{L-1} + alpha p_{L-1}           r_L = r_{L-1} - alpha q_{L-1}
//   rho_L = rho_{L-1} - 2 alpha z.q + alpha^2 q.M^-1 q   (r_L.z_L expanded:
//                                            valid for any A, no symmetry used)
//   beta = rho_L / rho_{L-1}                p_L = z_L + beta p_{L-1}
//   q_L = A p_L -> the partials of iteration L
// LASpack's stop test |r| < eps |b| (RTCResult rtc.c:48-72, at the top of
// every iteration) on r_{L-1} runs at the start of launch L on the directly
// summed r_{L-1}.r_{L-1}; x_{L-1} of a stopped solve was written by launch
// L-1.  The IsOne(alpha) / IsOne(beta) shortcuts of LASpack's scaled vector
// operations are kept (a factor within 10 DBL_EPSILON of 1 is taken as 1).
// Launch 0 tests r_0 (partials r.r, r.z, b.b of the assembly, or of
// k_pcg1_start from a nonzero start) and forms p_0 = z_0 and the partials of
// q_0.  One launch reads r, p, x, D, cX, cY and writes r, p, x: 72 B per
// cell.  r and p alternate between two buffers (r_L in r / q by L & 1, p_L
// in pBuf[L & 1]), because the halo lanes read their neighbours' previous
// values; x is read and written only at the thread's own cells.
//
// Tile walk as in k_cg1: W adjacent columns per lane (W = 1, 2: one double
// or one 16-byte access per array and row), loads two rows ahead, row t
// ingested (q_{L-1}, r_L, p_L of row t) while row t-1 is finished (q_L, the
// stores, the partials); lanes 0 and 63 also evaluate the column v beyond
// the wave from the column o after it, with the owner's expressions.
namespace {

typedef double vd2 __attribute__((ext_vector_type(2)));

__device__ inline Rsrc rsrc_of(const double *p, int n)
{
    const uint64_t v = (uint64_t)p;
    const uint64_t u = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return __builtin_amdgcn_make_buffer_rsrc((void *)u, (short)0, __builtin_amdgcn_readfirstlane(n * 8), 0x00020000);
}

__device__ inline double dload1(Rsrc r, uint32_t off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ inline void dload(Rsrc r, uint32_t off, double (&o)[1]) { o[0] = dload1(r, off); }
__device__ inline void dload(Rsrc r, uint32_t off, double (&o)[2])
{
    const vd2 v = __builtin_bit_cast(vd2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    o[0] = v.x;
    o[1] = v.y;
}

// byte offset of (row, col) in an N-double vector, or OFF_NONE outside the
// grid or when !ok (N <= 2^28, checked at create)
__device__ inline uint32_t doff_of(int row, int col, int nx, int ny, bool ok)
{
    return (ok && row >= 0 && row < ny) ? (uint32_t)(row * nx + col) * 8u : OFF_NONE;
}

template <int W>
__device__ inline void dload_cols(Rsrc p, int row, int i, int nx, int ny, bool ok, double (&o)[W])
{
    dload(p, doff_of(row, i, nx, ny, ok), o);
}

__device__ inline void dstore_cols(double *p, size_t c, const double (&v)[1]) { gst(p + c, v[0]); }
__device__ inline void dstore_cols(double *p, size_t c, const double (&v)[2])
{
    gst(reinterpret_cast<vd2 *>(p + c), vd2{v[0], v[1]});
}

// (A p) of one cell, Mul_QV's row order (operats.c:870-888, columns sorted:
// west, south, centre, north, east); D < 0: an identity row scaled by -D.
// Neighbours outside the grid come in as zeros.
__device__ inline double apply_row(double pc, double Dc, double cW, double pW, double cS, double pS, double cN,
                                   double pN, double cE, double pE)
{
#pragma clang fp contract(off)
    if (Dc < 0) return -Dc * pc;
    double s = 0.0;
    s += (-cW) * pW;
    s += (-cS) * pS;
    s += Dc * pc;
    s += (-cN) * pN;
    s += (-cE) * pE;
    return s;
}

// r_it and p_it of a cell from r_{it-1}, q_{it-1}, p_{it-1} (the same
// rounding in the owner's lane and in a halo lane)
__device__ inline void advance_jacobi(double R, double q, double P, double inv, double al, double be, double &rn,
                                      double &pn)
{
#pragma clang fp contract(off)
    rn = R - al * q;
    pn = rn * inv + be * P;
}

// Jacobi's 1 / |D|, and 0 where D is 0: lanes past the grid's last column
// and the row past its last load zeros, and their p must come out 0 (their
// neighbours multiply it by a face coefficient, which is not 0 on the north
// edge)
__device__ inline double inv_diag(double D) { return D != 0.0 ? 1.0 / fabs(D) : 0.0; }

template <int W>
struct Pcg1Row {
    double P2[W], R[W], D[W], cX[W], cY[W], X[W];   // own columns: p(t+1), r(t), D(t), cX(t), cY(t), x(t-1)
    double hP2, hR, hD, hcXv, hcY, hPo, hcXo;       // lanes 0 / 63: column v and its outer neighbour o
};

}  // namespace

template <bool First, int W>
__global__ void __launch_bounds__(256) k_pcg1(const PcgArgs<double> *__restrict__ A, int it)
{
    it = First ? 0 : it;
    const PcgArgs<double> &a = A[blockIdx.z];
    if ((int)blockIdx.x >= a.nbDir) return;
    __shared__ double lds[16 * 5];
    PcgScalars *sc = a.sc;
    const int halted = *a.halted, doneAt = sc->doneAt;
    if (halted || (doneAt >= 0 && doneAt < it)) return;

    double alphaD = 0.0, betaD = 0.0;
    bool stop;
    if (First) {
        double v[3];
        reduce_partials<3>(a.partA, a.nbStart, v, lds);
        const double rNorm = sqrt(v[0]), bNorm = sqrt(v[2]);
        stop = rNorm < a.eps * bNorm || (LAS_IS_ZERO(bNorm) && LAS_IS_ONE(1.0 + rNorm)) || a.maxIter <= 0;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->bb = v[2];
            sc->bNorm = bNorm;
            sc->rr = v[0];
            sc->rNorm = rNorm;
            sc->iter = 0;
            if (stop) sc->doneAt = 0;
        }
    } else {
        double v[5];
        reduce_partials<5>(a.partF[(it - 1) & 1], a.nbDir, v, lds);
        const double pq = v[0], zq = v[1], qMq = v[2], rr = v[3], rz = v[4];
        const double rNorm = sqrt(rr), bNorm = sc->bNorm;
        stop = (it >= 2 && (rNorm < a.eps * bNorm || (LAS_IS_ZERO(bNorm) && LAS_IS_ONE(1.0 + rNorm)))) ||
               it - 1 >= a.maxIter || !(pq > 0);
        alphaD = rz / pq;
        if (LAS_IS_ONE(alphaD)) alphaD = 1.0;
        double rho = rz - 2.0 * alphaD * zq + alphaD * alphaD * qMq;
        if (!(rho > 0)) rho = 0;               // cancellation at convergence: restart the directions
        betaD = rho / rz;
        if (LAS_IS_ONE(betaD)) betaD = 1.0;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->rr = rr;
            sc->rNorm = rNorm;
            sc->iter = stop ? it - 1 : it;
            if (stop) {
                sc->doneAt = it;
            } else {
                sc->pq = pq;
                sc->alpha = alphaD;
                sc->rhoHist[(it - 1) & 1] = rz;
                sc->rhoHist[it & 1] = rho;
            }
        }
    }
    if (stop) return;                          // x_{it-1} is final
    const double al = alphaD, be = betaD;

    int lb = blockIdx.x;
    if (a.remap) lb = (lb & 7) * (a.nbDir >> 3) + (lb >> 3);
    const int bx = lb % a.gx, by = lb / a.gx;
    double part[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (by < a.gy) {
        const int nx = a.nx, ny = a.ny;
        const int lane = threadIdx.x & 63;
        const int i0 = bx * 256 * W + (threadIdx.x >> 6) * 64 * W;   // the wave's first column
        const int i = i0 + lane * W;
        const bool valid = i < nx;                                   // nx is a multiple of W
        const int j0 = by * a.H, j1 = min(j0 + a.H, ny);
        double *const rBuf[2] = {a.r, a.q};
        const double *Rin = rBuf[First ? 0 : (it - 1) & 1];
        const double *Pin = a.pBuf[(it - 1) & 1];                    // unused at it = 0
        double *Rout = rBuf[it & 1];
        double *Pout = a.pBuf[it & 1];
        double *X = a.x;
        const Rsrc cXr = rsrc_of(a.cX, a.N), cYr = rsrc_of(a.cY, a.N), Dr = rsrc_of(a.D, a.N);
        const Rsrc rIn = rsrc_of(Rin, a.N), pIn = rsrc_of(Pin, a.N), xR = rsrc_of(X, a.N);
        // halo lanes: v = the column beyond the wave, o = the one after it
        const int v = lane == 0 ? i0 - 1 : (lane == 63 ? i0 + 64 * W : -1);
        const bool halo = v >= 0 && v < nx;
        const int o = lane == 0 ? v - 1 : v + 1;
        const bool oin = halo && o >= 0 && o < nx;
        auto at = [&](Rsrc p, int row, int col, bool ok) -> double {
            return dload1(p, doff_of(row, col, nx, ny, ok));
        };
        auto load = [&](int t, Pcg1Row<W> &L) {
            dload_cols<W>(pIn, t + 1, i, nx, ny, valid && !First, L.P2);
            dload_cols<W>(rIn, t, i, nx, ny, valid, L.R);
            dload_cols<W>(Dr, t, i, nx, ny, valid, L.D);
            dload_cols<W>(cXr, t, i, nx, ny, valid, L.cX);
            dload_cols<W>(cYr, t, i, nx, ny, valid, L.cY);
            dload_cols<W>(xR, t - 1, i, nx, ny, valid && !First && t - 1 >= j0, L.X);
            L.hP2 = at(pIn, t + 1, v, halo && !First);
            L.hR = at(rIn, t, v, halo);
            L.hD = at(Dr, t, v, halo);
            L.hcXv = at(cXr, t, v, halo);
            L.hcY = at(cYr, t, v, halo);
            L.hPo = at(pIn, t, o, oin && !First);
            L.hcXo = at(cXr, t, o, oin && lane == 0 && !First);
        };
        // window at the start: rows j0-2 and j0-1 of p_{it-1}, the south
        // face of row j0-1 and the diagonal of row j0-1 come from the first
        // ingested row
        double P0[W], P1[W], cYm[W], cYmm[W], cXc[W], Dm[W], pnm[W], pnc[W], rnc[W];
        dload_cols<W>(pIn, j0 - 2, i, nx, ny, valid && !First, P0);
        dload_cols<W>(pIn, j0 - 1, i, nx, ny, valid && !First, P1);
        dload_cols<W>(cYr, j0 - 2, i, nx, ny, valid, cYm);
#pragma unroll
        for (int k = 0; k < W; k++) cYmm[k] = cXc[k] = Dm[k] = pnm[k] = pnc[k] = rnc[k] = 0.0;
        double hP0 = at(pIn, j0 - 2, v, halo && !First), hP1 = at(pIn, j0 - 1, v, halo && !First);
        double hcYm = at(cYr, j0 - 2, v, halo), hpn = 0.0, hcVIprev = 0.0;
        Pcg1Row<W> b0, b1, b2;
        load(j0 - 1, b0);
        load(j0, b1);
        auto row = [&](int t, const Pcg1Row<W> &cur, Pcg1Row<W> &nxt2) {
            if (t + 2 <= j1) load(t + 2, nxt2);
            // own columns: q_{it-1}, r_it, p_it of row t
            double pWl = __shfl_up(P1[W - 1], 1, 64), pEl = __shfl_down(P1[0], 1, 64);
            double cWl = __shfl_up(cur.cX[W - 1], 1, 64);
            if (lane == 0) { pWl = hP1; cWl = cur.hcXv; }
            if (lane == 63) pEl = hP1;
            double rn[W], pn[W];
#pragma unroll
            for (int k = 0; k < W; k++) {
                const double inv = inv_diag(cur.D[k]);
                if (First) {
                    rn[k] = cur.R[k];
                    pn[k] = cur.R[k] * inv;
                } else {
                    const double pW = k == 0 ? pWl : P1[k - 1];
                    const double pE = k == W - 1 ? pEl : P1[k + 1];
                    const double cW = k == 0 ? cWl : cur.cX[k - 1];
                    const double q = apply_row(P1[k], cur.D[k], cW, pW, cYm[k], P0[k], cur.cY[k], cur.P2[k],
                                               cur.cX[k], pE);
                    advance_jacobi(cur.R[k], q, P1[k], inv, al, be, rn[k], pn[k]);
                }
            }
            // halo column v of row t, the owner's expressions, evaluated in
            // every lane and kept by lanes 0 and 63
            double hpnT;
            {
                const double inv = inv_diag(cur.hD);
                if (First) {
                    hpnT = cur.hR * inv;
                } else {
                    const double hq0 = apply_row(hP1, cur.hD, cur.hcXo, cur.hPo, hcYm, hP0, cur.hcY, cur.hP2,
                                                 cur.hcXv, P1[0]);
                    const double hq63 = apply_row(hP1, cur.hD, cur.cX[W - 1], P1[W - 1], hcYm, hP0, cur.hcY,
                                                  cur.hP2, cur.hcXv, cur.hPo);
                    double hr;
                    advance_jacobi(cur.hR, lane == 0 ? hq0 : hq63, hP1, inv, al, be, hr, hpnT);
                }
            }
            hpnT = halo ? hpnT : 0.0;
            // finish row j = t-1: q_it = A p_it, stores, partials
            const int j = t - 1;
            double qWl = __shfl_up(pnc[W - 1], 1, 64), qEl = __shfl_down(pnc[0], 1, 64);
            double cWjl = __shfl_up(cXc[W - 1], 1, 64);
            if (lane == 0) { qWl = hpn; cWjl = hcVIprev; }
            if (lane == 63) qEl = hpn;
            if (valid && j >= j0) {
                const size_t c = (size_t)j * nx + i;
                double xs[W];
#pragma unroll
                for (int k = 0; k < W; k++) {
                    const double qW = k == 0 ? qWl : pnc[k - 1];
                    const double qE = k == W - 1 ? qEl : pnc[k + 1];
                    const double cWj = k == 0 ? cWjl : cXc[k - 1];
                    const double qn = apply_row(pnc[k], Dm[k], cWj, qW, cYmm[k], pnm[k], cYm[k], pn[k], cXc[k], qE);
                    const double inv = 1.0 / fabs(Dm[k]);
                    const double zc = rnc[k] * inv;
                    xs[k] = cur.X[k] + al * P0[k];
                    part[0] += pnc[k] * qn;
                    part[1] += zc * qn;
                    part[2] += qn * (qn * inv);
                    part[3] += rnc[k] * rnc[k];
                    part[4] += rnc[k] * zc;
                }
                if (!First) {
                    dstore_cols(Rout, c, rnc);
                    dstore_cols(X, c, xs);
                }
                dstore_cols(Pout, c, pnc);
            }
            // slide
#pragma unroll
            for (int k = 0; k < W; k++) {
                P0[k] = P1[k]; P1[k] = cur.P2[k];
                cYmm[k] = cYm[k]; cYm[k] = cur.cY[k];
                cXc[k] = cur.cX[k];
                Dm[k] = cur.D[k];
                pnm[k] = pnc[k]; pnc[k] = pn[k]; rnc[k] = rn[k];
            }
            hP0 = hP1; hP1 = cur.hP2;
            hcYm = cur.hcY;
            hpn = hpnT;
            hcVIprev = cur.hcXv;
        };
        for (int t = j0 - 1; t <= j1; t += 3) {
            row(t, b0, b2);
            if (t + 1 > j1) break;
            row(t + 1, b1, b0);
            if (t + 2 > j1) break;
            row(t + 2, b2, b1);
        }
    }
    block_sum<5>(part, lds);
    store_partials<5>(a.partF[it & 1], a.nbDir, {part[0], part[1], part[2], part[3], part[4]});
}

// The start residual from a nonzero start x_0 (CGIter itersolv.c:388-393:
// r = b - A x_0), one cell per thread over a grid-stride loop, with the
// partials r.r, r.z and b.b for k_pcg1's launch 0.
__global__ void __launch_bounds__(256) k_pcg1_start(const PcgArgs<double> *__restrict__ A)
{
    const PcgArgs<double> &a = A[blockIdx.z];
    if ((int)blockIdx.x >= a.nbStart || !a.warm) return;
    __shared__ double lds[16 * 3];
    if (*a.halted) return;
    const double *b = a.b;
    double part[3] = {0.0, 0.0, 0.0};
    const int nx = a.nx, ny = a.ny;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < a.N; c += a.nbStart * blockDim.x) {
        const int i = c % nx, j = c / nx;
        const double Dc = a.D[c];
        const double ax = apply_row(a.x[c], Dc, i > 0 ? a.cX[c - 1] : 0.0, i > 0 ? a.x[c - 1] : 0.0,
                                    j > 0 ? a.cY[c - nx] : 0.0, j > 0 ? a.x[c - nx] : 0.0, a.cY[c],
                                    j < ny - 1 ? a.x[c + nx] : 0.0, a.cX[c], i < nx - 1 ? a.x[c + 1] : 0.0);
        const double bc = b[c];
        const double r = bc - ax;
        a.r[c] = r;
        part[0] += r * r;
        part[1] += r * (r * (1.0 / fabs(Dc)));
        part[2] += bc * bc;
    }
    block_sum<3>(part, lds);
    store_partials<3>(a.partA, a.nbStart, {part[0], part[1], part[2]});
}

// ---------------------------------------------------------------------------
// The fused mixed-precision solve: every launch of the solve (k_refine,
// k_cg1, the flag gather, k_surface) as rounds of one persistent kernel.
//
// Rounds and barriers.  In a round each block does the work of one launch
// of the launched form for its member -- refinement sweep s, inner iteration
// it, or the surface update -- over its share of the member's tiles, each
// tile's block-reduced sums stored in the partial slot the launched form's
// block for that tile uses.  A grid-wide barrier follows; then every block
// takes the decision the next launch would take at its start (the outer
// test of sweep s, the inner stop test with alpha and beta of the next
// iteration) from the same partials in the same fixed order, so every block
// of a member obtains bitwise the same decision and the results are those of
// the launched form, bit for bit.  The member's block 0 records the solver
// scalars and, once the solve has converged, writes the member's flags.
// Nothing is launched past a member's convergence, and the host does not
// take part in the solve.
//
// The barrier: each block's thread 0, after an agent-scope release fence of
// every thread (the block's stores written back from its XCD's L2), reads
// the generation, adds its vote (work in this round) and its arrival; the
// last arrival resets the arrivals and the next round's votes and advances
// the generation; the others wait for the generation to move (an acquire
// load, so the round's stores of every block are visible after it).  A wait
// longer than spinTicks of the device's wall clock sets the timeout flag,
// and every block leaves at its next barrier: a launch whose blocks were not
// all resident ends with the solve unfinished instead of waiting forever.
// The launch leaves the barrier state as it found it (zero arrivals and
// votes; the generation is only compared).
namespace {

enum { PH_REFINE, PH_CG, PH_SURF, PH_DONE };

// returns the votes of this round (0: no block had work), or -1 after a timeout
__device__ inline int grid_barrier(SolveSync *sy, unsigned nblocks, int round, int vote,
                                   unsigned long long spinTicks, int *sh)
{
    // the block's stores are complete in its XCD's L2 after the workgroup
    // barrier; thread 0's agent-scope release writes them back, its acquire
    // after the wait invalidates the stale lines once for the block (the
    // waiting itself polls with coherent relaxed loads, which leave the
    // caches alone)
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned g = __hip_atomic_load(&sy->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (vote) __hip_atomic_fetch_add(&sy->votes[round & 1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned arrived = __hip_atomic_fetch_add(&sy->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int err = 0;
        if (arrived == nblocks - 1) {
            // every block's stores happen before this release of the
            // generation: acquire the arrivals first
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(&sy->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&sy->votes[(round + 1) & 1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&sy->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const unsigned long long t0 = wall_clock64();
            while (__hip_atomic_load(&sy->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (__hip_atomic_load(&sy->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    err = 1;
                    break;
                }
                if ((unsigned long long)wall_clock64() - t0 > spinTicks) {
                    __hip_atomic_store(&sy->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    err = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        sh[0] = err ? -1 : (int)__hip_atomic_load(&sy->votes[round & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return sh[0];
}

// block `index` of `count`'s share [lo, hi) of n items
__device__ inline void share(int n, int index, int count, int &lo, int &hi)
{
    lo = (int)((long long)n * index / count);
    hi = (int)((long long)n * (index + 1) / count);
}

}  // namespace

template <int W>
__global__ void __launch_bounds__(256) k_solve_mixed(const RefineArgs *__restrict__ R,
                                                     const PcgArgs<float> *__restrict__ P,
                                                     const StepArgsT<float> *__restrict__ S,
                                                     const SolveBlock *__restrict__ plan, SolveSync *sync,
                                                     int32_t *flagsOut, unsigned long long spinTicks)
{
    const SolveBlock pb = plan[blockIdx.x];
    const int m = pb.member;
    const RefineArgs &ra = R[m];
    const PcgArgs<float> &pa = P[m];
    const StepArgsT<float> &sa = S[m];
    __shared__ double lds[16 * 4];
    __shared__ int sh[2];
    Ctl *ctl = ra.ctl;
    const bool rec = pb.index == 0;        // the member's recording block
    const bool halted = ctl->halted != 0;
    if (rec && threadIdx.x == 0) {
        // until the surface is written, the solve is open (a launch that
        // ends at a timeout leaves it so, and the kernels after the solve
        // stand down)
        ctl->open = halted ? 0 : 1;
        if (halted) {
            close_solve(ctl, true, true);
            write_flags(ctl, flagsOut + (size_t)m * F_STRIDE);
        }
    }
    int ph = halted ? PH_DONE : PH_REFINE;
    int s = 0, it = 0;
    double eps = 0.0, bNorm = 0.0;
    float al = 0.0f, be = 0.0f;
    bool stop0 = false;                     // the inner solve stopped at its set-up
    int lo, hi;
    for (int round = 0;; round++) {
        if (ph == PH_REFINE) {
            share(ra.nb, pb.index, pb.count, lo, hi);
            for (int lb = lo; lb < hi; lb++) {
                double part[4] = {0.0, 0.0, 0.0, 0.0};
                refine_tile(ra, s, lb, part);
                block_sum<4>(part, lds);
                refine_store(ra, slot_of(lb, ra.nb, ra.remap), part);
            }
        } else if (ph == PH_CG) {
            share(pa.nbDir, pb.index, pb.count, lo, hi);
            for (int lb = lo; lb < hi; lb++) {
                double part[4] = {0.0, 0.0, 0.0, 0.0};
                if (it == 0) cg1_tile<true, W>(pa, it, al, be, stop0, lb, part);
                else cg1_tile<false, W>(pa, it, al, be, stop0, lb, part);
                block_sum<4>(part, lds);
                if (!stop0 && threadIdx.x == 0) {
                    const int slot = slot_of(lb, pa.nbDir, pa.remap);
                    double *pf = pa.partF[it & 1];
#pragma unroll
                    for (int q = 0; q < 4; q++) pf[(size_t)q * pa.nbDir + slot] = part[q];
                }
            }
        } else if (ph == PH_SURF) {
            // the surface from the converged sweep's x (x_s in xbuf[s & 1]),
            // in the chunks of 256 indices of k_surface, whose partials it
            // writes
            const gptr<const double> xs = (s & 1) ? sa.xB : sa.x;
            const int nb1 = (sa.g.Ng + 255) / 256;
            share(nb1, pb.index, pb.count, lo, hi);
            for (int ch = lo; ch < hi; ch++) {
                double loss[1] = {0.0};
                const int k = ch * 256 + (int)threadIdx.x;
                if (k < sa.g.Ng) surface_cell(sa, k, loss[0], xs);
                block_sum<1>(loss, lds);
                if (threadIdx.x == 0) sa.vpart[2 * sa.maxBlocks + ch] += loss[0];
            }
            if (rec && threadIdx.x == 0) {
                close_solve(ctl, true, true);
                write_flags(ctl, flagsOut + (size_t)m * F_STRIDE);
            }
        }
        const int total = grid_barrier(sync, gridDim.x, round, ph == PH_REFINE || ph == PH_CG, spinTicks, sh);
        if (total <= 0) break;
        // the decision the next launch would take at its start
        if (ph == PH_REFINE) {
            const RefTest t = refine_test(ra, s, lds, rec);
            if (t.conv) {
                ph = PH_SURF;
            } else {
                ph = PH_CG;
                it = 0;
                eps = t.epsIn;
                const Cg1Dec d = cg1_decide(pa, 0, eps, 0.0, lds, rec);
                stop0 = d.stop;
                bNorm = d.rNorm0;
                al = be = 0.0f;
            }
        } else if (ph == PH_CG) {
            bool ended = it == 0 && stop0;
            if (!ended) {
                const Cg1Dec d = cg1_decide(pa, it + 1, eps, bNorm, lds, rec);
                if (d.stop) {
                    ended = true;
                } else {
                    it += 1;
                    al = (float)d.alphaD;
                    be = (float)d.betaD;
                }
            }
            if (ended) {
                ph = PH_REFINE;
                s += 1;
                stop0 = false;
            }
        } else if (ph == PH_SURF) {
            ph = PH_DONE;
        }
    }
}

template __global__ void k_pcg_dir<double>(const PcgArgs<double> *__restrict__, int);
template __global__ void k_pcg_res<double>(const PcgArgs<double> *__restrict__, int);
template __global__ void k_ssor<true>(const PcgArgs<double> *__restrict__, int);
template __global__ void k_ssor<false>(const PcgArgs<double> *__restrict__, int);
template __global__ void k_cg1<true, 1>(const PcgArgs<float> *__restrict__, int, const RefineArgs *__restrict__, int);
template __global__ void k_cg1<false, 1>(const PcgArgs<float> *__restrict__, int, const RefineArgs *__restrict__, int);
template __global__ void k_cg1<true, 2>(const PcgArgs<float> *__restrict__, int, const RefineArgs *__restrict__, int);
template __global__ void k_cg1<false, 2>(const PcgArgs<float> *__restrict__, int, const RefineArgs *__restrict__, int);
template __global__ void k_cg1<true, 4>(const PcgArgs<float> *__restrict__, int, const RefineArgs *__restrict__, int);
template __global__ void k_cg1<false, 4>(const PcgArgs<float> *__restrict__, int, const RefineArgs *__restrict__, int);
template __global__ void k_pcg1<true, 1>(const PcgArgs<double> *__restrict__, int);
template __global__ void k_pcg1<false, 1>(const PcgArgs<double> *__restrict__, int);
template __global__ void k_pcg1<true, 2>(const PcgArgs<double> *__restrict__, int);
template __global__ void k_pcg1<false, 2>(const PcgArgs<double> *__restrict__, int);
#define FREHD_SOLVE_MIXED(W)                                                                             \
    template __global__ void k_solve_mixed<W>(const RefineArgs *__restrict__, const PcgArgs<float> *__restrict__, \
                                              const StepArgsT<float> *__restrict__, const SolveBlock *__restrict__, \
                                              SolveSync *, int32_t *, unsigned long long);
FREHD_SOLVE_MIXED(1)
FREHD_SOLVE_MIXED(2)
FREHD_SOLVE_MIXED(4)

}  // namespace frehd
