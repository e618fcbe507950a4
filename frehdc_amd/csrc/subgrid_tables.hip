// subgrid_tables.hip -- the subgrid level tables generated on the device from
// a fine bathymetry of r x r pixels per coarse cell (the values of each
// level: cell_levels in subgrid_eval.h).  Tables are written interleaved,
// tabs[(k*N + c)*10 + T] (tab_index, the device layout of the step kernels),
// c = j*nx + i.  One thread per coarse cell walks a chunk of levels; its
// pixels stay in L1/L2 across levels.
//
// k_tables_reorder moves one table between the reference's level-major
// layout t[k*N + c] (files, host arrays; FP64) and the interleaved one.
// The interleaved tables are stored in the handle's state precision TT
// (FP32 for the mixed form: values rounded to nearest).
#include <hip/hip_runtime.h>

#include "step_kernels.h"
#include "subgrid_eval.h"

namespace frehd {

template <class TT>
struct TableGenArgs {
    const double *fine;       // fy x fx, row-major (x fastest), fx = nx*r
    TT *tabs;                 // interleaved V Z Vxp Vxm Vyp Vym Np Nm Op Om
    double surfmin, dsurf, dxf;
    int nx, ny, r, Ne;
    int k0, k1;               // level range of this launch
};

template <class TT>
__global__ void __launch_bounds__(256) k_subgrid_tables(TableGenArgs<TT> a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const FineBed bed{(gptr<const double>)a.fine, a.nx, a.ny, a.r, a.dxf};
    const size_t c = (size_t)j * a.nx + i;
    for (int k = a.k0; k < a.k1; k++) {
        const double h[1] = {level_raw(a.surfmin, a.dsurf, k)};
        double v[1][10];
        cell_levels<1>(bed, i, j, h, v);
        TT *o = a.tabs + tab_index(c, k, (size_t)a.nx * a.ny, 0);
#pragma unroll
        for (int t = 0; t < 10; t++) o[t] = (TT)v[0][t];
    }
}

// one table between level-major lm[k*N + c] and the interleaved layout;
// toInterleaved selects the direction
template <class TT>
__global__ void __launch_bounds__(256) k_tables_reorder(TT *tabs, double *lm, int N, int Ne, int T,
                                                        int toInterleaved)
{
    const size_t total = (size_t)N * Ne;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
        // q = k*N + c on both sides
        const int k = (int)(q / N);
        const size_t c = q - (size_t)k * N;
        if (toInterleaved) tabs[tab_index(c, k, N, T)] = (TT)lm[q];
        else lm[q] = tabs[tab_index(c, k, N, T)];
    }
}

int launch_tables_reorder(void *tabs, double *lm, int N, int Ne, int T, int toInterleaved, int f32, hipStream_t stream)
{
    if (f32)
        hipLaunchKernelGGL(k_tables_reorder<float>, dim3(4096), dim3(256), 0, stream, static_cast<float *>(tabs), lm, N,
                           Ne, T, toInterleaved);
    else
        hipLaunchKernelGGL(k_tables_reorder<double>, dim3(4096), dim3(256), 0, stream, static_cast<double *>(tabs), lm,
                           N, Ne, T, toInterleaved);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// host launcher: levels in chunks so that no launch runs long
template <class TT>
static int launch_tables(const double *fineDev, TT *tabs, int nx, int ny, int r, int Ne, double surfmin, double dsurf,
                         double dxf, hipStream_t stream)
{
    TableGenArgs<TT> a;
    a.fine = fineDev;
    a.tabs = tabs;
    a.surfmin = surfmin;
    a.dsurf = dsurf;
    a.dxf = dxf;
    a.nx = nx;
    a.ny = ny;
    a.r = r;
    a.Ne = Ne;
    const dim3 b(64, 4), g((nx + 63) / 64, (ny + 3) / 4);
    const int chunk = 16;
    for (int k0 = 0; k0 < Ne; k0 += chunk) {
        a.k0 = k0;
        a.k1 = k0 + chunk < Ne ? k0 + chunk : Ne;
        hipLaunchKernelGGL(k_subgrid_tables<TT>, g, b, 0, stream, a);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

int launch_subgrid_tables(const double *fineDev, void *tabs, int nx, int ny, int r, int Ne, double surfmin,
                          double dsurf, double dxf, int f32, hipStream_t stream)
{
    if (f32) return launch_tables(fineDev, static_cast<float *>(tabs), nx, ny, r, Ne, surfmin, dsurf, dxf, stream);
    return launch_tables(fineDev, static_cast<double *>(tabs), nx, ny, r, Ne, surfmin, dsurf, dxf, stream);
}

}  // namespace frehd
