// subgrid_gen.cpp -- the subgrid level tables from a fine bathymetry, on the
// host (the pre-processing of the reference, preProcessing/
// script_CreateSubAreaWithBlockV16.m:144-460 with ComputeFaceAreaWithBlockCheck.m
// or ComputeSubgridFaceArea.m), and their text files.
//
// Levels are independent, so they are spread over host threads; within a
// level every sweep keeps the script's loop order where a later write can
// replace an earlier one (the minimum-volume pass writes neighbours'
// entries).  Sums run in the script's order: a whole cell as one
// column-major vector, a half cell or a face neighbourhood over the pixels
// of a column (the first index) first, then over the columns.
//
// The script's quirks that decide values are kept and marked below:
// the x-direction minimum areas sum unmasked depths while the y-direction
// ones mask the pixels the cleaning removed; the minimum areas carry no
// pixel width (the face areas do); the cross-cell connectivity pass starts
// at the second row and the second column; the last row's minimum-volume
// pass writes the first row's Vxm; face areas are taken before the blocked
// faces are zeroed by the minimum-area pass and after it by the
// minimum-volume pass.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "frehd_gpu.h"
#include "frehd_subgrid.h"

namespace frehd {
int fail(int code, const char *fmt, ...);
}
using frehd::fail;

namespace {

// MATLAB round (half away from zero) of a non-negative x
inline int mround(double x) { return (int)floor(x + 0.5); }

struct Ctx {
    const double *fine;
    int fnx, fny, r, nx, ny, N, Ne;
    double dx, dA, Dx, surfmin, dsurf, n0;
    int checkBlock, useminV, useYh, useCn;
    std::vector<double> bottom;
    double bath(int X, int Y) const { return fine[(size_t)Y * fnx + X]; }
};

enum { T_V, T_Z, T_VXP, T_VXM, T_VYP, T_VYM, T_NP, T_NM, T_OP, T_OM, T_NMIN, T_NMAX, T_OMIN, T_OMAX,
       T_MVXP, T_MVXM, T_MVYP, T_MVYM, T_YHX, T_YHY, T_CN, T_COUNT };

const char *const kNames[T_COUNT] = {"V", "Z", "Vxp", "Vxm", "Vyp", "Vym", "Np", "Nm", "Op", "Om", "Nmin",
                                     "Nmax", "Omin", "Omax", "minVxp", "minVxm", "minVyp", "minVym", "YhX",
                                     "YhY", "Cn"};

double **out_slot(frehd_subgrid_out *o, int t)
{
    double **s[T_COUNT] = {&o->V, &o->Z, &o->Vxp, &o->Vxm, &o->Vyp, &o->Vym, &o->Np, &o->Nm, &o->Op, &o->Om,
                           &o->Nmin, &o->Nmax, &o->Omin, &o->Omax, &o->minVxp, &o->minVxm, &o->minVyp,
                           &o->minVym, &o->YhX, &o->YhY, &o->Cn};
    return s[t];
}

// one thread's working memory for a level
struct Work {
    std::vector<double> tab;        // T_COUNT * N: a level's tables
    std::vector<uint8_t> clean;     // fine wet flags after the per-cell cleaning
    std::vector<uint8_t> block;     // 4 per cell: Np, Nm, Op, Om blocked
    std::vector<int> label, queue;  // r * r
    std::vector<int> size;
    std::vector<double> sums, sorted, sumsEx;
};

// Steps 1 and 2 of the block check (ComputeFaceAreaWithBlockCheck.m:28-77):
// keep the largest 4-connected wet cluster of the cell (clusters numbered in
// the column-major order of their first pixel, as bwlabel numbers them; the
// first of equal largest wins), none when it holds under 1 % of the pixels;
// a face with no wet pixel left on it is blocked.
void clean_cell(const Ctx &c, double h, int ii, int jj, Work &w)
{
    const int r = c.r, X0 = ii * r, Y0 = jj * r;
    int *lab = w.label.data();
    // a is the pixel's first (x) index in the cell, b the second; the
    // column-major position is a + r b
    for (int q = 0; q < r * r; q++) lab[q] = 0;
    w.size.assign(1, 0);
    int nl = 0;
    for (int b = 0; b < r; b++)
        for (int a = 0; a < r; a++) {
            const int p = a + r * b;
            if (lab[p] || !(h > c.bath(X0 + a, Y0 + b))) continue;
            nl++;
            w.size.push_back(0);
            int head = 0, tail = 0;
            w.queue[tail++] = p;
            lab[p] = nl;
            while (head < tail) {
                const int q = w.queue[head++];
                w.size[nl]++;
                const int qa = q % r, qb = q / r;
                const int na[4] = {qa - 1, qa + 1, qa, qa};
                const int nb[4] = {qb, qb, qb - 1, qb + 1};
                for (int e = 0; e < 4; e++) {
                    if (na[e] < 0 || na[e] >= r || nb[e] < 0 || nb[e] >= r) continue;
                    const int t = na[e] + r * nb[e];
                    if (lab[t] || !(h > c.bath(X0 + na[e], Y0 + nb[e]))) continue;
                    lab[t] = nl;
                    w.queue[tail++] = t;
                }
            }
        }
    int keep = 0;
    if (nl > 0) {
        keep = 1;
        for (int l = 2; l <= nl; l++)
            if (w.size[l] > w.size[keep]) keep = l;
        if (w.size[keep] < 0.01 * r * r) keep = 0;      // small water area removed
    }
    for (int b = 0; b < r; b++)
        for (int a = 0; a < r; a++)
            w.clean[(size_t)(Y0 + b) * c.fnx + X0 + a] = keep && lab[a + r * b] == keep;
    bool wNm = false, wNp = false, wOm = false, wOp = false;
    for (int q = 0; q < r; q++) {
        wNm = wNm || w.clean[(size_t)(Y0 + q) * c.fnx + X0];
        wNp = wNp || w.clean[(size_t)(Y0 + q) * c.fnx + X0 + r - 1];
        wOm = wOm || w.clean[(size_t)Y0 * c.fnx + X0 + q];
        wOp = wOp || w.clean[(size_t)(Y0 + r - 1) * c.fnx + X0 + q];
    }
    uint8_t *bk = &w.block[4 * (size_t)(jj * c.nx + ii)];
    bk[0] = !wNp;
    bk[1] = !wNm;
    bk[2] = !wOp;
    bk[3] = !wOm;
}

double median_of(std::vector<double> &v)
{
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return n % 2 ? v[n / 2] : (v[n / 2 - 1] + v[n / 2]) / 2;
}

// Step 5 along one axis (ComputeFaceAreaWithBlockCheck.m:209-322): the
// per-row sums of depth across the face's neighbourhood, their minimum,
// and the face's maximum area (its area, or the minimum where an internal
// contraction is detected).  `sums`: the r (or, at the domain edge, r/2)
// row sums; `ex`: the wider neighbourhood's sums (interior faces only).
void min_max_area(const Ctx &c, Work &w, bool interior, double mid, bool nbOk, double &amin, double &amax,
                  bool &set)
{
    const std::vector<double> &s = w.sums;
    w.sorted.clear();
    for (double v : s)
        if (v != 0) w.sorted.push_back(v);
    set = !w.sorted.empty();
    if (!set) return;
    const double mn = *std::min_element(w.sorted.begin(), w.sorted.end());
    const double mx = *std::max_element(w.sorted.begin(), w.sorted.end());
    const size_t len = w.sorted.size();
    const double med = median_of(w.sorted);
    amin = mn;
    amax = mid;
    if (!interior) return;
    const double mnEx = *std::min_element(w.sumsEx.begin(), w.sumsEx.end());
    const bool ccFirst = w.sumsEx.front() == mnEx, ccLast = w.sumsEx.back() == mnEx;
    if (s.front() != 0 && s.back() != 0 && nbOk && (double)len > 0.5 * c.r && !ccFirst && !ccLast && len > 3) {
        if (fabs(mx - med) > 2 * fabs(med - mn)) amax = mn;
        else if (fabs(med - mn) > 2 * fabs(mx - med)) amax = mn;
    }
}

void level(const Ctx &c, int k, Work &w, frehd_subgrid_out *out)
{
    const int r = c.r, nx = c.nx, ny = c.ny, N = c.N;
    const double h = c.surfmin + k * c.dsurf;
    double *T[T_COUNT];
    for (int t = 0; t < T_COUNT; t++) {
        double *o = *out_slot(out, t);
        T[t] = o ? o + (size_t)k * N : &w.tab[(size_t)t * N];
        std::fill(T[t], T[t] + N, 0.0);
    }
    auto depth = [&](int X, int Y) { return std::max(h - c.bath(X, Y), 0.0); };
    auto cl = [&](int X, int Y) { return w.clean[(size_t)Y * c.fnx + X] != 0; };
    const int h2 = mround(r / 2.0);

    // face areas: with the block check (steps 1-4), or plain
    if (c.checkBlock && r > 1) {
        for (int jj = 0; jj < ny; jj++)
            for (int ii = 0; ii < nx; ii++) clean_cell(c, h, ii, jj, w);
        // step 3: wet pixels facing each other across the face, for cells
        // from the second row and column on (the script's loops start at 2)
        for (int ii = 1; ii < nx; ii++)
            for (int jj = 1; jj < ny; jj++) {
                const int X = ii * r, Y = jj * r;
                bool fx = false, fy = false;
                for (int q = 0; q < r; q++) {
                    fx = fx || (cl(X, Y + q) && cl(X - 1, Y + q));
                    fy = fy || (cl(X + q, Y) && cl(X + q, Y - 1));
                }
                if (!fx) {
                    w.block[4 * (size_t)(jj * nx + ii) + 1] = 1;
                    w.block[4 * (size_t)(jj * nx + ii - 1) + 0] = 1;
                }
                if (!fy) {
                    w.block[4 * (size_t)(jj * nx + ii) + 3] = 1;
                    w.block[4 * (size_t)((jj - 1) * nx + ii) + 2] = 1;
                }
            }
    }
    const bool masked = c.checkBlock && r > 1;
    auto dm = [&](int X, int Y) { return masked && !cl(X, Y) ? 0.0 : depth(X, Y); };
    if (r > 1) {
        // step 4 (or ComputeSubgridFaceArea.m:8-38): the mutually wet depth
        // of the two pixel columns at the face; at the domain's last face the
        // cell's own column, and the first cell's m face from the first
        for (int ii = 0; ii < nx; ii++)
            for (int jj = 0; jj < ny; jj++) {
                const int X2 = ii * r + r - 1, Y2 = jj * r + r - 1, c0 = jj * nx + ii;
                double s = 0.0;
                if (ii < nx - 1) {
                    for (int q = 0; q < r; q++) s += std::min(dm(X2, jj * r + q), dm(X2 + 1, jj * r + q));
                    T[T_NP][c0] = s * c.dx;
                    T[T_NM][c0 + 1] = s * c.dx;
                } else {
                    for (int q = 0; q < r; q++) s += depth(X2, jj * r + q);
                    T[T_NP][c0] = s * c.dx;
                    double s0 = 0.0;
                    for (int q = 0; q < r; q++) s0 += depth(0, jj * r + q);
                    T[T_NM][jj * nx] = s0 * c.dx;
                }
                s = 0.0;
                if (jj < ny - 1) {
                    for (int q = 0; q < r; q++) s += std::min(dm(ii * r + q, Y2), dm(ii * r + q, Y2 + 1));
                    T[T_OP][c0] = s * c.dx;
                    T[T_OM][c0 + nx] = s * c.dx;
                } else {
                    for (int q = 0; q < r; q++) s += depth(ii * r + q, Y2);
                    T[T_OP][c0] = s * c.dx;
                    double s0 = 0.0;
                    for (int q = 0; q < r; q++) s0 += depth(ii * r + q, 0);
                    T[T_OM][ii] = s0 * c.dx;
                }
            }
    }
    if (masked) {
        // step 5: minimum and maximum face areas (from the face areas before
        // the blocked faces are zeroed)
        const int lo = r / 2, hi = h2;                               // floor(r/2), round(r/2)
        const int loEx = (3 * r) / 4, hiEx = mround(3.0 * r / 4.0);   // floor(3r/4), round(3r/4)
        for (int ii = 0; ii < nx; ii++)
            for (int jj = 0; jj < ny; jj++) {
                const int X2 = ii * r + r - 1, Y2 = jj * r + r - 1, c0 = jj * nx + ii;
                bool set;
                double amin = 0, amax = 0;
                // x: rows across the face, unmasked (the script's mask is
                // commented out on this axis)
                const bool inX = ii < nx - 1;
                w.sums.clear();
                for (int X = X2 - lo + 1; X <= (inX ? X2 + hi : X2); X++) {
                    double s = 0.0;
                    for (int q = 0; q < r; q++) s += depth(X, jj * r + q);
                    w.sums.push_back(s);
                }
                w.sumsEx.clear();
                if (inX)
                    for (int X = X2 - loEx + 1; X <= X2 + hiEx; X++) {
                        double s = 0.0;
                        for (int q = 0; q < r; q++) s += depth(X, jj * r + q);
                        w.sumsEx.push_back(s);
                    }
                const bool nbX = inX && T[T_NP][c0 + 1] > 0 && T[T_NM][c0] > 0;
                min_max_area(c, w, inX, T[T_NP][c0], nbX, amin, amax, set);
                if (set) {
                    T[T_NMIN][c0] = amin;
                    T[T_NMAX][c0] = amax;
                }
                // y: columns across the face, the cleaned-away pixels masked
                const bool inY = jj < ny - 1;
                w.sums.clear();
                for (int Y = Y2 - lo + 1; Y <= (inY ? Y2 + hi : Y2); Y++) {
                    double s = 0.0;
                    for (int q = 0; q < r; q++) s += dm(ii * r + q, Y);
                    w.sums.push_back(s);
                }
                w.sumsEx.clear();
                if (inY)
                    for (int Y = Y2 - loEx + 1; Y <= Y2 + hiEx; Y++) {
                        double s = 0.0;
                        for (int q = 0; q < r; q++) s += depth(ii * r + q, Y);
                        w.sumsEx.push_back(s);
                    }
                const bool nbY = inY && T[T_OP][c0 + nx] > 0 && T[T_OM][c0] > 0;
                min_max_area(c, w, inY, T[T_OP][c0], nbY, amin, amax, set);
                if (set) {
                    T[T_OMIN][c0] = amin;
                    T[T_OMAX][c0] = amax;
                }
            }
    }
    // volumes, wet area, the blocked faces zeroed, drag-model tables
    // (script V16 :222-340)
    for (int ii = 0; ii < nx; ii++)
        for (int jj = 0; jj < ny; jj++) {
            const int X0 = ii * r, Y0 = jj * r, c0 = jj * nx + ii;
            // the cell's pixels as one column-major vector (sum(localDepth(:)))
            double sv = 0.0, wet = 0.0;
            for (int b = 0; b < r; b++)
                for (int a = 0; a < r; a++) {
                    const double d = depth(X0 + a, Y0 + b);
                    sv += d;
                    wet += d > 0 ? 1.0 : 0.0;
                }
            T[T_V][c0] = c.dA * sv;
            T[T_Z][c0] = c.dA * wet;
            if (r > 1) {
                // half volumes: sum over the first index, then the second
                auto half = [&](int a0, int a1, int b0, int b1) {
                    double s = 0.0;
                    for (int b = b0; b < b1; b++) {
                        double col = 0.0;
                        for (int a = a0; a < a1; a++) col += depth(X0 + a, Y0 + b);
                        s += col;
                    }
                    return c.dA * s;
                };
                T[T_VXP][c0] = half(h2, r, 0, r);
                T[T_VYP][c0] = half(0, r, h2, r);
                T[T_VXM][c0] = half(0, h2, 0, r);
                T[T_VYM][c0] = half(0, r, 0, h2);
                if (masked) {
                    const uint8_t *bk = &w.block[4 * (size_t)c0];
                    if (bk[0]) T[T_NP][c0] = 0;
                    if (bk[1]) T[T_NM][c0] = 0;
                    if (bk[2]) T[T_OP][c0] = 0;
                    if (bk[3]) T[T_OM][c0] = 0;
                }
            } else {
                const double d = depth(X0, Y0);
                T[T_VXP][c0] = T[T_VYP][c0] = T[T_VXM][c0] = T[T_VYM][c0] = c.dA * d / 2;
                T[T_NP][c0] = T[T_OP][c0] = T[T_NM][c0] = T[T_OM][c0] = c.dx * d;
            }
            if (c.useYh) {
                const double Y = T[T_V][c0] / (c.Dx * c.Dx);
                double sh = 0.0;
                bool any = false;
                for (int b = 0; b < r; b++)
                    for (int a = 0; a < r; a++) {
                        const double d = depth(X0 + a, Y0 + b);
                        if (d > 0) {
                            sh += pow(d, 5.0 / 3.0);
                            any = true;
                        }
                    }
                const double v = any ? (c.Dx * c.Dx / (c.dx * c.dx)) * pow(Y, 3.0) * pow(sh, -2.0) : 0.0;
                T[T_YHX][c0] = v;
                T[T_YHY][c0] = v;
            }
            if (c.useCn) {
                // Casas 2010: the mean height of the submerged relief over the
                // cell bottom, emergent pixels at the bottom
                const double zb = c.bottom[c0];
                double m = 0.0;
                for (int b = 0; b < r; b++) {
                    double col = 0.0;
                    for (int a = 0; a < r; a++) {
                        const double z = c.bath(X0 + a, Y0 + b);
                        col += (z > h ? zb : z) - zb;
                    }
                    m += col / r;
                }
                const double D = std::max(0.0, m / r), H = std::max(0.0, h - zb);
                const double xi = D > 0 ? H / D : 100.0;
                if (xi >= 0.2 && xi <= 7) {
                    const double fxi = 1 + (1 / xi) * log(cosh(1 - xi) / cosh(1.0));
                    T[T_CN][c0] = pow(H, 1.0 / 6.0) / sqrt(9.81 * 4.5 * fxi);
                } else {
                    T[T_CN][c0] = c.n0;
                }
            }
        }
    if (!c.useminV) return;
    // minimum half-cell volumes (script V16 :342-440), in the script's loop
    // order: a cell's pass writes its neighbour's m entry, which a later
    // cell's pass may replace
    const int a0 = h2;
    for (int ii = 0; ii < nx; ii++)
        for (int jj = 0; jj < ny; jj++) {
            const int X0 = ii * r, Y0 = jj * r, c0 = jj * nx + ii;
            // x: the rows from the cell's middle to the next cell's middle
            {
                const int Xa = X0 + a0, Xb = ii < nx - 1 ? X0 + r - 1 + a0 : X0 + r - 1;
                double mn = 0.0;
                bool any = false;
                int wetPix = 0;
                for (int X = Xa; X <= Xb; X++) {
                    double s = 0.0;
                    for (int q = 0; q < r; q++) {
                        const double d = depth(X, Y0 + q);
                        s += d;
                        wetPix += d > 0;
                    }
                    if (s != 0) {
                        mn = any ? std::min(mn, s) : s;
                        any = true;
                    }
                }
                if (any) {
                    const double minV = c.Dx * (mn * c.dx);
                    if (minV > 0) {
                        if (T[T_NMAX][c0] < T[T_NP][c0]) {
                            T[T_MVXP][c0] = 0.5 * minV;
                            if (ii != nx - 1) T[T_MVXM][c0 + 1] = 0.5 * minV;
                            else T[T_MVXM][jj * nx] = T[T_MVXP][jj * nx];
                        } else if ((double)wetPix / (r * r) > 0.2) {
                            T[T_MVXP][c0] = std::max(std::min(0.5 * minV, T[T_VXP][c0]), 0.7 * T[T_VXP][c0]);
                            if (ii != nx - 1)
                                T[T_MVXM][c0 + 1] =
                                    std::max(std::min(0.5 * minV, T[T_VXM][c0 + 1]), 0.7 * T[T_VXM][c0 + 1]);
                            else
                                T[T_MVXM][jj * nx] = T[T_MVXP][jj * nx];
                        } else {
                            T[T_MVXP][c0] = T[T_VXP][c0];
                            T[T_MVXM][c0] = T[T_VXM][c0];
                        }
                    }
                } else {
                    T[T_MVXP][c0] = T[T_VXP][c0];
                    T[T_MVXM][c0] = T[T_VXM][c0];
                }
            }
            // y: the columns from the cell's middle to the next cell's middle
            {
                const int Ya = Y0 + a0, Yb = jj < ny - 1 ? Y0 + r - 1 + a0 : Y0 + r - 1;
                double mn = 0.0;
                bool any = false;
                int wetPix = 0;
                for (int Y = Ya; Y <= Yb; Y++) {
                    double s = 0.0;
                    for (int q = 0; q < r; q++) {
                        const double d = depth(X0 + q, Y);
                        s += d;
                        wetPix += d > 0;
                    }
                    if (s != 0) {
                        mn = any ? std::min(mn, s) : s;
                        any = true;
                    }
                }
                if (any) {
                    const double minV = c.Dx * (mn * c.dx);
                    if (minV > 0) {
                        if (T[T_OMAX][c0] < T[T_OP][c0]) {
                            T[T_MVYP][c0] = 0.5 * minV;
                            if (jj != ny - 1) T[T_MVYM][c0 + nx] = 0.5 * minV;
                            else T[T_MVYM][ii] = T[T_MVYP][ii];
                        } else if ((double)wetPix / (r * r) > 0.2) {
                            T[T_MVYP][c0] = std::max(std::min(0.5 * minV, T[T_VYP][c0]), 0.7 * T[T_VYP][c0]);
                            if (jj != ny - 1)
                                T[T_MVYM][c0 + nx] =
                                    std::max(std::min(0.5 * minV, T[T_VYM][c0 + nx]), 0.7 * T[T_VYM][c0 + nx]);
                            else
                                T[T_MVYM][ii] = T[T_MVYP][ii];
                        } else {
                            T[T_MVYP][c0] = T[T_VYP][c0];
                            T[T_MVYM][c0] = T[T_VYM][c0];
                        }
                    }
                } else {
                    T[T_MVYP][c0] = T[T_VYP][c0];
                    T[T_MVYM][c0] = T[T_VYM][c0];
                }
            }
        }
}

int make_dirs(const std::string &path)
{
    std::string cur;
    for (size_t q = 0; q <= path.size(); q++) {
        if (q == path.size() || path[q] == '/') {
            if (!cur.empty() && cur != "/") {
                struct stat st;
                if (stat(cur.c_str(), &st) != 0 && mkdir(cur.c_str(), 0755) != 0)
                    return fail(FREHD_EINVAL, "cannot create folder %s", cur.c_str());
            }
        }
        if (q < path.size()) cur.push_back(path[q]);
    }
    return 0;
}

int check_spec(const frehd_subgrid_spec *s)
{
    if (!s) return fail(FREHD_EINVAL, "spec is NULL");
    if (s->r < 1 || s->r > 256) return fail(FREHD_EINVAL, "r %d outside [1, 256]", s->r);
    if (s->nxFine % s->r || s->nyFine % s->r)
        return fail(FREHD_EINVAL, "the fine grid (%d x %d) is not a whole number of %d x %d cells", s->nxFine,
                    s->nyFine, s->r, s->r);
    if (s->nxFine / s->r < 2 || s->nyFine / s->r < 2) return fail(FREHD_EINVAL, "fewer than 2 x 2 coarse cells");
    if (s->checkBlock && s->r > 1 && s->r % 2)
        return fail(FREHD_EINVAL, "the block check splits cells into halves: r must be even (r = %d)", s->r);
    if (s->Ne < 1) return fail(FREHD_EINVAL, "Ne %d < 1", s->Ne);
    if (!(s->dxf > 0) || !(s->dsurf > 0)) return fail(FREHD_EINVAL, "dxf and dsurf must be positive");
    return 0;
}

}  // namespace

extern "C" int frehd_subgrid_generate(const double *fine, const frehd_subgrid_spec *spec, frehd_subgrid_out *out,
                                      int32_t threads)
{
    if (int rc = check_spec(spec)) return rc;
    if (!fine || !out) return fail(FREHD_EINVAL, "NULL argument");
    Ctx c;
    c.fine = fine;
    c.fnx = spec->nxFine;
    c.fny = spec->nyFine;
    c.r = spec->r;
    c.nx = c.fnx / c.r;
    c.ny = c.fny / c.r;
    c.N = c.nx * c.ny;
    c.Ne = spec->Ne;
    c.dx = spec->dxf;
    c.dA = c.dx * c.dx;
    c.Dx = c.r * c.dx;
    c.surfmin = spec->surfmin;
    c.dsurf = spec->dsurf;
    c.n0 = spec->n0;
    c.checkBlock = spec->checkBlock != 0;
    c.useminV = spec->useminVolume != 0;
    c.useYh = spec->useYh != 0;
    c.useCn = spec->useCn != 0;
    // cell bottom (the lowest pixel) and face bottoms (the higher of the two
    // lowest pixels of the pixel columns at the face; at the domain edge the
    // cell's own), script V16 :152-172
    c.bottom.resize(c.N);
    for (int jj = 0; jj < c.ny; jj++)
        for (int ii = 0; ii < c.nx; ii++) {
            const int X0 = ii * c.r, Y0 = jj * c.r, c0 = jj * c.nx + ii;
            double mn = c.bath(X0, Y0);
            for (int b = 0; b < c.r; b++)
                for (int a = 0; a < c.r; a++) mn = std::min(mn, c.bath(X0 + a, Y0 + b));
            c.bottom[c0] = mn;
            auto colmin = [&](int X) {
                double m = c.bath(X, Y0);
                for (int q = 1; q < c.r; q++) m = std::min(m, c.bath(X, Y0 + q));
                return m;
            };
            auto rowmin = [&](int Y) {
                double m = c.bath(X0, Y);
                for (int q = 1; q < c.r; q++) m = std::min(m, c.bath(X0 + q, Y));
                return m;
            };
            const int X2 = X0 + c.r - 1, Y2 = Y0 + c.r - 1;
            const double bx = ii < c.nx - 1 ? std::max(colmin(X2), colmin(X2 + 1)) : colmin(X2);
            const double by = jj < c.ny - 1 ? std::max(rowmin(Y2), rowmin(Y2 + 1)) : rowmin(Y2);
            if (out->bottom) out->bottom[c0] = mn;
            if (out->bottomXP) out->bottomXP[c0] = bx;
            if (out->bottomYP) out->bottomYP[c0] = by;
        }
    int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    nt = std::max(1, std::min(nt, c.Ne));
    auto work = [&](int t) {
        Work w;
        w.tab.resize((size_t)T_COUNT * c.N);
        w.clean.assign((size_t)c.fnx * c.fny, 0);
        w.block.assign(4 * (size_t)c.N, 0);
        w.label.resize((size_t)c.r * c.r);
        w.queue.resize((size_t)c.r * c.r);
        for (int k = t; k < c.Ne; k += nt) {
            std::fill(w.block.begin(), w.block.end(), 0);
            level(c, k, w, out);
        }
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> ts;
        for (int t = 0; t < nt; t++) ts.emplace_back(work, t);
        for (auto &t : ts) t.join();
    }
    return 0;
}

extern "C" int frehd_subgrid_write_text(const char *folder, const frehd_subgrid_spec *spec,
                                        const frehd_subgrid_out *t)
{
    if (int rc = check_spec(spec)) return rc;
    if (!folder || !t) return fail(FREHD_EINVAL, "NULL argument");
    if (int rc = make_dirs(folder)) return rc;
    const size_t N = (size_t)(spec->nxFine / spec->r) * (spec->nyFine / spec->r);
    frehd_subgrid_out copy = *t;
    struct Item {
        const char *name;
        const double *v;
        size_t n;
    };
    std::vector<Item> items;
    for (int q = 0; q < T_COUNT; q++)
        if (*out_slot(&copy, q)) items.push_back(Item{kNames[q], *out_slot(&copy, q), N * (size_t)spec->Ne});
    if (t->bottom) items.push_back(Item{"bottom", t->bottom, N});
    if (t->bottomXP) items.push_back(Item{"bottomXP", t->bottomXP, N});
    if (t->bottomYP) items.push_back(Item{"bottomYP", t->bottomYP, N});
    for (const Item &it : items) {
        std::string path = std::string(folder) + "/" + it.name + ".dat";
        FILE *f = fopen(path.c_str(), "w");
        if (!f) return fail(FREHD_EINVAL, "cannot write %s", path.c_str());
        for (size_t q = 0; q < it.n; q++) fprintf(f, "%f\n", it.v[q]);
        if (fclose(f) != 0) return fail(FREHD_EINVAL, "write error on %s", path.c_str());
    }
    return 0;
}
