// subgrid_eval.h -- the ten subgrid values of one coarse cell at given
// levels, from its r x r fine pixels (and the adjacent pixel column / row of
// its neighbours for the face areas).
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
//
// Both users -- the table generator (every level) and the on-the-fly lookup
// of the step (the two levels bracketing the surface) -- call cell_levels,
// with floating-point contraction off, so a value computed on the fly is
// bitwise the table value: the order of the sums does not depend on how many
// levels one call evaluates.
#pragma once

#include <hip/hip_runtime.h>

#include "frehd_dev.h"

namespace frehd {

// raw elevation of level k, two roundings (no fused multiply-add)
__host__ __device__ inline double level_raw(double surfmin, double dsurf, int k)
{
#pragma clang fp contract(off)
    const double m = (double)k * dsurf;
    return surfmin + m;
}

struct FineBed {
    gptr<const double> z;   // (ny*r) rows of (nx*r) raw elevations, x fastest
    int nx, ny, r;
    double dxf;
};

// values of cell (i, j) at the NL raw levels h[0..NL-1]: out[l][T] in the
// table order V Z Vxp Vxm Vyp Vym Np Nm Op Om.  RR > 0: the pixel count per
// cell side known at compile time (= b.r), so the pixel loops unroll and
// every pixel load of the cell is in flight at once; the sums keep their
// order, so the values are bitwise those of RR = 0.
template <int NL, int RR = 0>
__device__ inline void cell_levels(const FineBed &b, int i, int j, const double (&h)[NL], double (&out)[NL][10])
{
#pragma clang fp contract(off)
    const int r = RR > 0 ? RR : b.r, h2 = (r + 1) / 2;
    const size_t fx = (size_t)b.nx * r;
    const double dA = b.dxf * b.dxf;
    const double *z0 = b.z + (size_t)j * r * fx + (size_t)i * r;   // pixel (x=0, y=0) of the cell
    double vol[NL], vxp[NL], vxm[NL], vyp[NL], vym[NL], wet[NL];
    double np[NL], nm[NL], op[NL], om[NL];
#pragma unroll
    for (int l = 0; l < NL; l++) {
        vol[l] = vxp[l] = vxm[l] = vyp[l] = vym[l] = wet[l] = 0.0;
        np[l] = nm[l] = op[l] = om[l] = 0.0;
    }
#pragma unroll
    for (int y = 0; y < r; y++) {
        const double *row = z0 + (size_t)y * fx;
#pragma unroll
        for (int x = 0; x < r; x++) {
            const double zz = row[x];
#pragma unroll
            for (int l = 0; l < NL; l++) {
                const double d = fmax(h[l] - zz, 0.0);
                vol[l] += d;
                if (d > 0) wet[l] += 1.0;
                if (x >= h2) vxp[l] += d; else vxm[l] += d;
                if (y >= h2) vyp[l] += d; else vym[l] += d;
            }
        }
    }
    if (r > 1) {
        const bool east = i < b.nx - 1, west = i > 0, north = j < b.ny - 1, south = j > 0;
#pragma unroll
        for (int y = 0; y < r; y++) {
            const double *row = z0 + (size_t)y * fx;
            const double zl = row[r - 1], zf = row[0];
            const double ze = east ? row[r] : 0.0, zw = west ? row[-1] : 0.0;
#pragma unroll
            for (int l = 0; l < NL; l++) {
                const double last = fmax(h[l] - zl, 0.0);
                np[l] += east ? fmin(last, fmax(h[l] - ze, 0.0)) : last;
                const double first = fmax(h[l] - zf, 0.0);
                nm[l] += west ? fmin(fmax(h[l] - zw, 0.0), first) : first;
            }
        }
        const double *top = z0 + (size_t)(r - 1) * fx;
#pragma unroll
        for (int x = 0; x < r; x++) {
            const double zl = top[x], zf = z0[x];
            const double zn = north ? top[x + fx] : 0.0, zs = south ? z0[(ptrdiff_t)x - (ptrdiff_t)fx] : 0.0;
#pragma unroll
            for (int l = 0; l < NL; l++) {
                const double last = fmax(h[l] - zl, 0.0);
                op[l] += north ? fmin(last, fmax(h[l] - zn, 0.0)) : last;
                const double first = fmax(h[l] - zf, 0.0);
                om[l] += south ? fmin(fmax(h[l] - zs, 0.0), first) : first;
            }
        }
    }
#pragma unroll
    for (int l = 0; l < NL; l++) {
        double *o = out[l];
        o[0] = dA * vol[l];
        o[1] = dA * wet[l];
        if (r > 1) {
            o[2] = dA * vxp[l];
            o[3] = dA * vxm[l];
            o[4] = dA * vyp[l];
            o[5] = dA * vym[l];
            o[6] = np[l] * b.dxf;
            o[7] = nm[l] * b.dxf;
            o[8] = op[l] * b.dxf;
            o[9] = om[l] * b.dxf;
        } else {
            const double half = dA * vol[l] / 2;
            for (int t = 2; t < 6; t++) o[t] = half;
            for (int t = 6; t < 10; t++) o[t] = b.dxf * vol[l];
        }
    }
}

}  // namespace frehd
