// subgrid_tables.hip -- the subgrid level tables generated on the device from
// a fine bathymetry of r x r pixels per coarse cell.
//
// Restates the reference's table generator (preProcessing/
// script_CreateSubAreaWithBlockV16.m:129-262 with checkBlock = 0; face areas
// as in ComputeSubgridFaceArea.m:1-43).  Level k is h = surfmin + k*dsurf
// (raw elevations, no offset); pixel depth D = max(h - z, 0), pixel area
// dxf^2.  For coarse cell (i, j), x = fine column within the cell, y = fine
// row, h2 = round(r/2):
//   V   = dxf^2 sum D            Z   = dxf^2 #{D > 0}
//   Vxp = dxf^2 sum_{x >= h2} D  Vxm = dxf^2 sum_{x < h2} D   (Vyp/Vym in y)
//   Np  = dxf sum_y min(D(last column), D(first column of cell i+1))
//         (the last column alone on the domain's east edge)
//   Nm  = Np of cell i-1 (the domain's first column alone at i = 0)
//   Op/Om the same across rows.  With r = 1 the halves are D dxf^2 / 2 and
//   the faces D dxf.
// Tables are written interleaved, tabs[(c*Ne + k)*10 + T] (the device
// layout of the step kernels), c = j*nx + i.  One thread per coarse cell
// walks all levels; its pixels stay in L1/L2 across levels.
//
// k_tables_reorder moves one table between the reference's level-major
// layout t[k*N + c] (files, host arrays) and the interleaved one.
#include <hip/hip_runtime.h>

#include "step_kernels.h"

namespace frehd {

struct TableGenArgs {
    const double *fine;       // fy x fx, row-major (x fastest), fx = nx*r
    double *tabs;             // interleaved V Z Vxp Vxm Vyp Vym Np Nm Op Om
    double surfmin, dsurf, dxf;
    int nx, ny, r, Ne;
    int k0, k1;               // level range of this launch
};

__global__ void __launch_bounds__(256) k_subgrid_tables(TableGenArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int r = a.r, fx = a.nx * r, h2 = (r + 1) / 2;
    const size_t c = (size_t)j * a.nx + i;
    const double dA = a.dxf * a.dxf;
    const double *z0 = a.fine + (size_t)j * r * fx + (size_t)i * r;   // pixel (x=0, y=0) of the cell
    for (int k = a.k0; k < a.k1; k++) {
        const double h = a.surfmin + k * a.dsurf;
        double vol = 0, vxp = 0, vxm = 0, vyp = 0, vym = 0, wet = 0;
        for (int y = 0; y < r; y++) {
            const double *row = z0 + (size_t)y * fx;
            for (int x = 0; x < r; x++) {
                const double d = fmax(h - row[x], 0.0);
                vol += d;
                if (d > 0) wet += 1.0;
                if (x >= h2) vxp += d; else vxm += d;
                if (y >= h2) vyp += d; else vym += d;
            }
        }
        double np = 0, nm = 0, op = 0, om = 0;
        if (r > 1) {
            for (int y = 0; y < r; y++) {
                const double *row = z0 + (size_t)y * fx;
                const double last = fmax(h - row[r - 1], 0.0);
                np += i < a.nx - 1 ? fmin(last, fmax(h - row[r], 0.0)) : last;
                const double first = fmax(h - row[0], 0.0);
                nm += i > 0 ? fmin(fmax(h - row[-1], 0.0), first) : first;
            }
            const double *top = z0 + (size_t)(r - 1) * fx;
            for (int x = 0; x < r; x++) {
                const double last = fmax(h - top[x], 0.0);
                op += j < a.ny - 1 ? fmin(last, fmax(h - top[x + fx], 0.0)) : last;
                const double first = fmax(h - z0[x], 0.0);
                om += j > 0 ? fmin(fmax(h - z0[x - fx], 0.0), first) : first;
            }
        }
        double *o = a.tabs + tab_index(c, k, a.Ne, 0);
        o[0] = dA * vol;
        o[1] = dA * wet;
        if (r > 1) {
            o[2] = dA * vxp;
            o[3] = dA * vxm;
            o[4] = dA * vyp;
            o[5] = dA * vym;
            o[6] = np * a.dxf;
            o[7] = nm * a.dxf;
            o[8] = op * a.dxf;
            o[9] = om * a.dxf;
        } else {
            const double half = dA * vol / 2;
            for (int t = 2; t < 6; t++) o[t] = half;
            for (int t = 6; t < 10; t++) o[t] = a.dxf * vol;
        }
    }
}

// one table between level-major lm[k*N + c] and the interleaved layout;
// toInterleaved selects the direction
__global__ void __launch_bounds__(256) k_tables_reorder(double *tabs, double *lm, int N, int Ne, int T,
                                                        int toInterleaved)
{
    const size_t total = (size_t)N * Ne;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
        // q walks the interleaved order (cell-major), so its stores are the
        // near-contiguous side
        const size_t c = q / Ne;
        const int k = (int)(q - c * Ne);
        const size_t l = (size_t)k * N + c;
        if (toInterleaved) tabs[tab_index(c, k, Ne, T)] = lm[l];
        else lm[l] = tabs[tab_index(c, k, Ne, T)];
    }
}

int launch_tables_reorder(double *tabs, double *lm, int N, int Ne, int T, int toInterleaved, hipStream_t stream)
{
    hipLaunchKernelGGL(k_tables_reorder, dim3(4096), dim3(256), 0, stream, tabs, lm, N, Ne, T, toInterleaved);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// host launcher: levels in chunks so that no launch runs long
int launch_subgrid_tables(const double *fineDev, double *tabs, int nx, int ny, int r, int Ne,
                          double surfmin, double dsurf, double dxf, hipStream_t stream)
{
    TableGenArgs a;
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
        hipLaunchKernelGGL(k_subgrid_tables, g, b, 0, stream, a);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

}  // namespace frehd
