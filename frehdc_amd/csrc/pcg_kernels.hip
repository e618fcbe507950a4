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
// With decouple set (the inner FP32 solve of the mixed path) a row ignores
// its neighbours that are identity rows, which makes the inner operator
// symmetric (block diagonal [I 0; 0 K]); the FP64 refinement around it
// computes the residual with the full operator, so the coupling enters
// through the next sweep's right-hand side.  Without it, CG's recurrences on
// the non-symmetric system stall in FP32 when the residual starts at the
// Dirichlet rows (a warm start right after the tide moves).
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

// raw values of one cell needed to form its new direction
template <typename T>
struct Raw {
    T r, D, po;
};

template <typename T>
__device__ inline Raw<T> load_raw(const PcgArgs<T> &a, const T *pOld, int c, int first)
{
    Raw<T> v;
    v.r = a.r[c];
    v.D = a.D[c];
    v.po = first ? T(0) : pOld[c];
    return v;
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
    const double rr = rs[0], rz = rs[1];
    const double rNorm = sqrt(rr);
    const double bNorm = it == 0 ? rNorm : sc->bNorm;
    const double eps = a.epsDyn ? *a.epsDyn : a.eps;
    const bool stop = rNorm < eps * bNorm || (LAS_IS_ZERO(bNorm) && LAS_IS_ONE(1.0 + rNorm)) ||
                      it >= a.maxIter;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (it == 0) { sc->bb = rr; sc->bNorm = bNorm; }
        sc->rr = rr;
        sc->rNorm = rNorm;
        sc->rhoHist[it & 1] = rz;
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
            T pS = 0, cyS = 0, DS = 1;
            if (valid && j0 > 0) {
                const int c = (j0 - 1) * nx + i;
                const Raw<T> v = load_raw(a, pOld, c, first);
                pS = direction(jacobi_z(v.r, v.D), v.po, beta, first, betaOne);
                cyS = a.cY[c];
                DS = v.D;
            }
            const bool dec = a.decouple != 0;
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
            T pc = direction(jacobi_z(cur.r, cur.D), cur.po, beta, first, betaOne);

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

                const T pn = (valid && j + 1 < ny) ? direction(jacobi_z(nxt.r, nxt.D), nxt.po, beta, first, betaOne)
                                                   : T(0);
                const T ph = halo ? direction(jacobi_z(hcur.r, hcur.D), hcur.po, beta, first, betaOne) : T(0);
                T pw = __shfl_up(pc, 1, 64), pe = __shfl_down(pc, 1, 64);
                T cxw = __shfl_up(cxc, 1, 64);
                T Dw = __shfl_up(cur.D, 1, 64), De = __shfl_down(cur.D, 1, 64);
                if (lane == 0) {
                    pw = ph;
                    Dw = hcur.D;
                    cxw = (valid && i > 0) ? a.cX[c - 1] : T(0);
                }
                if (lane == 63) {
                    pe = ph;
                    De = hcur.D;
                }
                if (valid) {
                    const T Dc = cur.D;
                    T sum;
                    if (Dc < 0) {
                        sum = -Dc * pc;
                    } else if (!dec) {
                        sum = 0;
                        if (i > 0) sum += (-cxw) * pw;
                        if (j > 0) sum += (-cyS) * pS;
                        sum += Dc * pc;
                        if (j < ny - 1) sum += (-cyc) * pn;
                        if (i < nx - 1) sum += (-cxc) * pe;
                    } else {
                        // symmetric inner operator: no coupling to identity rows
                        sum = 0;
                        if (i > 0 && !(Dw < 0)) sum += (-cxw) * pw;
                        if (j > 0 && !(DS < 0)) sum += (-cyS) * pS;
                        sum += Dc * pc;
                        if (j < ny - 1 && !(nxt.D < 0)) sum += (-cyc) * pn;
                        if (i < nx - 1 && !(De < 0)) sum += (-cxc) * pe;
                    }
                    pNew[c] = pc;
                    a.q[c] = sum;
                    part[0] += (double)pc * (double)sum;
                    a.x[c] = first ? T(0) : (alphaOne ? xc + cur.po : xc + al * cur.po);
                }
                pS = pc;
                cyS = cyc;
                DS = cur.D;
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
// k_refine(s): x_s = x_{s-1} + d (s > 0; x_0 is the start the assembly
// wrote) and the true residual r = b - A x_s with the FP64 operator, in the
// row order of the operator (W, S, C, N, E); the same sliding-window walk as
// k_pcg_dir, over x_s recomputed at the neighbours.  The inner solve runs on
// the decoupled operator (rows ignore identity-row neighbours), so its
// right-hand side is lifted: a row next to an identity row t gets
// + G * d_t, with d_t = r_t / (-D_t) the identity row's exact correction;
// the decoupled system then has the coupled system's solution.  Writes x_s,
// r32 = (float) (r + lift) and the partials b.b (s = 0), r.r (the FP64 true
// residual) and r32.r32, r32.z32 (the inner solve's first stop test and rho).
__global__ void __launch_bounds__(256) k_refine(const RefineArgs *__restrict__ A, int s)
{
    const RefineArgs &a = A[blockIdx.z];
    if ((int)blockIdx.x >= a.nb) return;
    __shared__ double lds[16 * 4];
    Ctl *ctl = a.ctl;
    if (ctl->halted || (ctl->ref.doneAt >= 0 && ctl->ref.doneAt < s)) return;
    const double *xin = s == 0 ? a.xbuf[0] : a.xbuf[(s + 1) & 1];   // x_0 or x_{s-1}
    double *xout = a.xbuf[s & 1];
    int lb = blockIdx.x;
    if (a.remap) lb = (lb & 7) * (a.nb >> 3) + (lb >> 3);
    const int bx = lb % a.gx, by = lb / a.gx;
    double part[4] = {0.0, 0.0, 0.0, 0.0};     // r32.r32, r32.z32, b.b, r.r
    if (by < a.gy) {
        const int nx = a.nx, ny = a.ny;
        const int lane = threadIdx.x & 63;
        const int i0 = bx * 256 + (threadIdx.x & ~63);
        const int i = i0 + lane;
        const bool valid = i < nx;
        const int j0 = by * a.H, j1 = min(j0 + a.H, ny);
        const int ih = lane == 0 ? i0 - 1 : (lane == 63 ? i0 + 64 : -1);
        const bool halo = ih >= 0 && ih < nx;
        auto xs = [&](int c) { return s > 0 ? xin[c] + (double)a.d[c] : xin[c]; };
        // rows j-1 (S), j (C), j+1 (N): x_s, diagonal and right-hand side
        double xS = 0, DS = 1, bS = 0, cyS = 0;
        if (valid && j0 > 0) {
            const int c = (j0 - 1) * nx + i;
            xS = xs(c);
            DS = a.D[c];
            bS = a.b[c];
            cyS = a.cY[c];
        }
        double xc = 0, Dc = 1, bc = 0, xn = 0, Dn = 1, bn = 0;
        if (valid) {
            const int c = j0 * nx + i;
            xc = xs(c);
            Dc = a.D[c];
            bc = a.b[c];
            if (j0 + 1 < ny) {
                xn = xs(c + nx);
                Dn = a.D[c + nx];
                bn = a.b[c + nx];
            }
        }
        for (int j = j0; j < j1; j++) {
            const int c = j * nx + i;
            double xh = 0, Dh = 1, bh = 0;
            if (halo) {
                const int h = j * nx + ih;
                xh = xs(h);
                Dh = a.D[h];
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
                xnn = xs(c + 2 * nx);
                Dnn = a.D[c + 2 * nx];
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
                const float r32 = (float)(r + lift);
                if (s > 0) xout[c] = xc;
                a.r32[c] = r32;
                part[0] += (double)r32 * (double)r32;
                part[1] += (double)r32 * (double)(r32 * (1.0f / fabsf(a.D32[c])));
                if (s == 0) part[2] += bc * bc;
                part[3] += r * r;
            }
            xS = xc; DS = Dc; bS = bc; cyS = cyc;
            xc = xn; Dc = Dn; bc = bn;
            xn = xnn; Dn = Dnn; bn = bnn;
        }
    }
    block_sum<4>(part, lds);
    if (threadIdx.x == 0) {
        a.partA[blockIdx.x] = part[0];
        a.partA[a.nb + blockIdx.x] = part[1];
        a.partC[blockIdx.x] = part[2];
        a.partC[a.nb + blockIdx.x] = part[3];
    }
}

// k_refine_test(s), one block: the outer stop test |b - A x_s| < eps |b|
// (the reference's criterion, rtc.c:54, on the FP64 residual) or s at the
// sweep cap; otherwise resets the inner solve with tolerance
// max(epsInner, eps |b| / (2 |r|)), so the last inner solve stops about when
// the outer test will pass.
__global__ void __launch_bounds__(256) k_refine_test(const RefineArgs *__restrict__ A, int s)
{
    const RefineArgs &a = A[blockIdx.z];
    __shared__ double lds[16 * 2];
    Ctl *ctl = a.ctl;
    if (ctl->halted || ctl->ref.doneAt >= 0) return;
    double v[2];
    reduce_partials<2>(a.partC, a.nb, v, lds);
    if (threadIdx.x != 0) return;
    RefineScalars &R = ctl->ref;
    if (s == 0) R.bNorm = sqrt(v[0]);
    const double bNorm = R.bNorm, rNorm = sqrt(v[1]);
    if (s > 0) {
        const int n = ctl->pcgIn.iter;
        if (s - 1 < 8) R.innerIters[s - 1] = n;
        ctl->pcg.iter += n;
    }
    R.rNorm = rNorm;
    R.sweep = s;
    const bool conv = rNorm < a.eps * bNorm || (LAS_IS_ZERO(bNorm) && LAS_IS_ONE(1.0 + rNorm)) ||
                      s >= a.maxRefine;
    if (conv) {
        R.doneAt = s;
        ctl->pcg.rNorm = rNorm;
        ctl->pcg.bNorm = bNorm;
        ctl->refineLast = s;
    } else {
        ctl->pcgIn.doneAt = -1;
        ctl->pcgIn.iter = 0;
        R.epsIn = fmax(a.epsInner, 0.5 * a.eps * bNorm / rNorm);
    }
}

template __global__ void k_pcg_dir<double>(const PcgArgs<double> *__restrict__, int);
template __global__ void k_pcg_res<double>(const PcgArgs<double> *__restrict__, int);
template __global__ void k_pcg_dir<float>(const PcgArgs<float> *__restrict__, int);
template __global__ void k_pcg_res<float>(const PcgArgs<float> *__restrict__, int);

}  // namespace frehd
