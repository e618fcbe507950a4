// frehd_dev.h -- device-side layout, index maps and reductions shared by the
// kernels of the free-surface step (gfx950, wave64).
//
// Layout: the reference's flat interior-first array with ghost blocks
// appended (FrehdC_v42/map.c:15-128), so x is the fastest index and every
// stencil read of an interior neighbour is a unit-stride (coalesced) access.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frehd {

// Pointer to device memory in the argument blocks.  The kernels read these
// pointers out of memory, so with plain pointers the compiler cannot tell
// global memory from LDS or scratch and emits flat instructions, and a flat
// access also counts against the wait counter of LDS and scalar loads: every
// scalar load of another pointer (and every cross-lane shuffle) would first
// drain all loads in flight.  Typed for the global address space in the
// device pass, the accesses become global loads and stores; the host pass
// sees plain pointers of the same size, and so does a source holding only
// host code (FREHD_HOST_SOURCE), which fills the blocks.
#if defined(__HIP_DEVICE_COMPILE__) && !defined(FREHD_HOST_SOURCE)
template <class T> using gptr = __attribute__((address_space(1))) T *;
#else
template <class T> using gptr = T *;
#endif

// An array of the explicit state stored in R and used in FP64.  In the device
// pass an FP32 array is a pointer wrapper whose elements read widened to
// double and store rounded to float, so every expression on the state
// computes in FP64 whatever the storage (the difference of two FP32 flow
// rates is formed in FP64, as the FP64 form forms it); an FP64 array is the
// plain global pointer.  On the host either is the plain pointer (same size
// and alignment), which is what the argument blocks carry.
#if defined(__HIP_DEVICE_COMPILE__) && !defined(FREHD_HOST_SOURCE)
template <class R>
struct WidePtr {
    gptr<R> p;
    struct Ref {
        gptr<R> q;
        __device__ operator double() const { return (double)*q; }
        __device__ const Ref &operator=(double v) const
        {
            *q = (R)v;
            return *this;
        }
        __device__ const Ref &operator=(const Ref &o) const
        {
            *q = *o.q;
            return *this;
        }
    };
    __device__ Ref operator[](size_t i) const { return Ref{p + i}; }
    __device__ WidePtr operator+(size_t off) const { return WidePtr{p + off}; }
};
template <class R> struct StatePtr { using type = WidePtr<R>; };
template <> struct StatePtr<double> { using type = gptr<double>; };
template <class R> using sptr = typename StatePtr<R>::type;
#else
template <class R> using sptr = R *;
#endif

struct Grid {
    int nx, ny, N, Ng;
};

// ghost block offsets (map.c:58-104)
__host__ __device__ inline int gjP(const Grid &g, int i) { return g.N + i; }
__host__ __device__ inline int gjM(const Grid &g, int i) { return g.N + g.nx + i; }
__host__ __device__ inline int giP(const Grid &g, int j) { return g.N + 2 * g.nx + j; }
__host__ __device__ inline int giM(const Grid &g, int j) { return g.N + 2 * g.nx + g.ny + j; }
__host__ __device__ inline int cIMJP(const Grid &g) { return g.Ng - 4; }
__host__ __device__ inline int cIPJP(const Grid &g) { return g.Ng - 3; }
__host__ __device__ inline int cIPJM(const Grid &g) { return g.Ng - 2; }
__host__ __device__ inline int cIMJM(const Grid &g) { return g.Ng - 1; }

// neighbours of interior cell c = j*nx + i
__host__ __device__ inline int nbE(const Grid &g, int c, int i, int j) { return i == g.nx - 1 ? giP(g, j) : c + 1; }
__host__ __device__ inline int nbW(const Grid &g, int c, int i, int j) { return i == 0 ? giM(g, j) : c - 1; }
__host__ __device__ inline int nbN(const Grid &g, int c, int i, int j) { return j == g.ny - 1 ? gjP(g, i) : c + g.nx; }
__host__ __device__ inline int nbS(const Grid &g, int c, int i, int j) { return j == 0 ? gjM(g, i) : c - g.nx; }
// two cells away (createMaps' iPPjc, iMMjc, icjPP, icjMM, map.c:68-98): the
// ghost beyond the edge where the second cell falls outside; icjMM of row 1,
// which map.c:69-74 never sets, is the jM ghost likewise
__host__ __device__ inline int nbEE(const Grid &g, int c, int i, int j) { return i >= g.nx - 2 ? giP(g, j) : c + 2; }
__host__ __device__ inline int nbWW(const Grid &g, int c, int i, int j) { return i <= 1 ? giM(g, j) : c - 2; }
__host__ __device__ inline int nbNN(const Grid &g, int c, int i, int j) { return j >= g.ny - 2 ? gjP(g, i) : c + 2 * g.nx; }
__host__ __device__ inline int nbSS(const Grid &g, int c, int i, int j) { return j <= 1 ? gjM(g, i) : c - 2 * g.nx; }

// Interior cell whose value a ghost (or corner) index mirrors under the
// reference's Neumann ghost fills (enforceFreeSurfaceBC nsfunctions.c:317-334,
// enforceVelocityBC :289-314 for its u corners, enforceScalarBC scalar.c:423-432).
// Returns -1 for an interior index.  `vcorners` selects the corner rule of
// vvYP/surf (corners copy iP/iM ghosts) vs uuXP (corners copy jP/jM ghosts).
__host__ __device__ inline int ghost_source(const Grid &g, int k, bool ucorners)
{
    const int nx = g.nx, ny = g.ny, N = g.N;
    if (k < N) return -1;
    int r = k - N;
    if (r < nx) return N - nx + r;                 // jP ghost <- top row
    r -= nx;
    if (r < nx) return r;                          // jM ghost <- bottom row
    r -= nx;
    if (r < ny) return r * nx + nx - 1;            // iP ghost <- right column
    r -= ny;
    if (r < ny) return r * nx;                     // iM ghost <- left column
    // corners
    if (ucorners) {
        if (k == cIPJP(g)) return N - 1;           // <- jPgt[nx-1] <- cell (nx-1, ny-1)
        if (k == cIPJM(g)) return nx - 1;          // <- jMgt[nx-1] <- cell (nx-1, 0)
        if (k == cIMJP(g)) return N - nx;          // <- jPgt[0]    <- cell (0, ny-1)
        return 0;                                  // iMjM <- jMgt[0] <- cell 0
    }
    if (k == cIPJP(g)) return (ny - 1) * nx + nx - 1;  // <- iPgt[ny-1]
    if (k == cIPJM(g)) return (ny - 1) * nx;           // <- iMgt[ny-1]
    if (k == cIMJP(g)) return nx - 1;                  // <- iPgt[0]
    return 0;                                          // iMjM <- iMgt[0]
}

// ---- reductions ------------------------------------------------------------
template <int NV>
struct Sums {
    double v[NV];
};

__device__ inline double wave_sum(double x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__device__ inline double wave_max(double x)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    return x;
}

// Block sum of NV values in a fixed order (deterministic for a given launch
// shape).  Result valid in thread 0.  blockDim.x*y*z must be a multiple of 64
// and at most 1024.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double *lds /* 16*NV */)
{
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6, nw = nthreads >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) lds[wid * NV + q] = v[q];
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double s = 0.0;
            for (int w = 0; w < nw; w++) s += lds[w * NV + q];
            v[q] = s;
        }
    }
    __syncthreads();
}

__device__ inline double block_max(double v, double *lds)
{
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6, nw = nthreads >> 6;
    v = wave_max(v);
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    if (tid == 0) {
        double m = lds[0];
        for (int w = 1; w < nw; w++) m = fmax(m, lds[w]);
        v = m;
    }
    __syncthreads();
    return v;
}

// ---- compensated (double-double) sums for the conservation budgets -------
// hi + lo carries a sum to about twice FP64's precision (Knuth's two-sum;
// the file is compiled without contraction, so no FMA alters the error terms)
struct DD {
    double hi, lo;
};

__host__ __device__ inline DD dd_two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}

__host__ __device__ inline DD dd_add(DD x, DD y)
{
    const DD s = dd_two_sum(x.hi, y.hi);
    const double e = s.lo + (x.lo + y.lo);
    const double h = s.hi + e;
    return DD{h, e - (h - s.hi)};
}

__host__ __device__ inline DD dd_add(DD x, double y) { return dd_add(x, DD{y, 0.0}); }

// block sum of a DD in a fixed order; valid in thread 0.  lds: 32 doubles.
__device__ inline DD block_dd_sum(DD v, double *lds)
{
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6, nw = nthreads >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const DD o{__shfl_xor(v.hi, off, 64), __shfl_xor(v.lo, off, 64)};
        v = dd_add(v, o);
    }
    if (lane == 0) {
        lds[2 * wid] = v.hi;
        lds[2 * wid + 1] = v.lo;
    }
    __syncthreads();
    if (tid == 0) {
        DD s{0.0, 0.0};
        for (int w = 0; w < nw; w++) s = dd_add(s, DD{lds[2 * w], lds[2 * w + 1]});
        v = s;
    }
    __syncthreads();
    return v;
}

// LASpack's IsZero / IsOne tests (elcmp.h:31-32): |a| < 10 DBL_MIN and
// |a - 1| < 10 DBL_EPSILON
#define LAS_IS_ZERO(a) (fabs(a) < 10.0 * 2.2250738585072014e-308)
#define LAS_IS_ONE(a) (fabs((a)-1.0) < 10.0 * 2.220446049250313e-16)

// ---- reductions without atomics ------------------------------------------
// Every reduction kernel runs a fixed grid of G blocks (grid-stride over the
// cells) and writes one partial per block.  The kernel that consumes the sum
// reduces all G partials itself, in every block, in the same fixed order, so
// every block obtains bitwise the same value and the result is identical run
// to run.  (Same-address atomics serialise across the chip, and agent-scope
// release/acquire fences write back and invalidate the per-XCD L2s; kernel
// boundaries give the visibility for free.)
template <int NV>
__device__ inline void store_partials(double *part, int nparts, const double (&v)[NV])
{
    // v already block-reduced into thread 0; nparts = the member's block count
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) part[(size_t)q * nparts + blockIdx.x] = v[q];
}

template <int NV>
__device__ inline void reduce_partials(const double *part, int nparts, double (&out)[NV], double *lds)
{
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) acc[q] = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
        for (int q = 0; q < NV; q++) acc[q] += part[(size_t)q * nparts + b];
    block_sum<NV>(acc, lds);
    __shared__ double bc[NV];
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) bc[q] = acc[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) out[q] = bc[q];
}

}  // namespace frehd
