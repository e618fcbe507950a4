// case_io.cpp -- FrehdC v4.2 input decks and output files (include/frehd_io.h).
//
// Host code only.  Each function cites the reference routine whose
// behaviour it restates; deviations (defined outcomes where the reference
// has undefined behaviour) are listed in frehd_io.h.
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include <string>
#include <vector>

#include "frehd_io.h"

namespace frehd {
int fail(int code, const char *fmt, ...);
}
using frehd::fail;

struct frehd_case {
    frehd_config cfg;
    frehd_case_info info;
    std::string inputFolder, outputFolder, subgridFolder;
    int useConstSurf0, useConstU0, useConstV0, useConstInitS, useminV, useCellEdge;
    double initSurf, initU, initV, initS;
    std::vector<int32_t> tideLocP, tideLocM, inflowLoc;
    std::vector<double> tideP, tideM, inflow, tidalPS, tidalMS, windspd, winddir;
    std::vector<double> allZ, allEdgeX, allEdgeY;
    std::vector<double> surf0, U0, V0, S0;
    std::vector<double> tab[14];   // V Z Vxp Vxm Vyp Vym Np Nm Op Om minVxp minVxm minVyp minVym
    std::vector<double> drag[3];   // YhX YhY (useSubDrag 2), Cn (3, 4)
    std::vector<double> mina[4];   // Nmin Nmax Omin Omax (useminA 1)
    frehd_subgrid sub;
};

namespace {

std::string join(const std::string &a, const std::string &b)
{
    if (a.empty()) return b;
    return a + b;   // the reference concatenates folder strings as given (they end in '/')
}

// ReadFile (fileio.c:27-41): N values read with fscanf("%lf,")
int read_values(const std::string &path, size_t N, std::vector<double> &out)
{
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return fail(FREHD_EINVAL, "cannot open %s", path.c_str());
    out.assign(N, 0.0);
    for (size_t q = 0; q < N; q++) {
        if (fscanf(f, "%lf,", &out[q]) != 1) {
            fclose(f);
            return fail(FREHD_EINVAL, "%s holds %zu values, %zu expected", path.c_str(), q, N);
        }
    }
    fclose(f);
    return 0;
}

// dateNum (utilities.c:98-110): mktime of year-month-day 00:00, no DST, in
// the process's local time zone
int date_num(const std::string &s, double *t)
{
    int y = 0, m = 0, d = 0;
    if (sscanf(s.c_str(), "%d-%d-%d", &y, &m, &d) != 3)
        return fail(FREHD_EINVAL, "start date '%s' is not YYYY-MM-DD", s.c_str());
    struct tm tmv;
    memset(&tmv, 0, sizeof tmv);
    tmv.tm_year = y - 1900;
    tmv.tm_mon = m - 1;
    tmv.tm_mday = d;
    tmv.tm_isdst = 0;
    *t = (double)mktime(&tmv);
    return 0;
}

// searchInd (utilities.c:129-150): first index whose time is within
// itval/2 (integer division, itval an int) of the target
int search_ind(const std::vector<double> &t, double targ, int itval, int N, int *ind)
{
    for (int q = 0; q < N; q++)
        if (fabs(t[q] - targ) <= itval / 2) {
            *ind = q;
            return 0;
        }
    return fail(FREHD_EINVAL, "no boundary-series time within %d s of %.0f", itval / 2, targ);
}

// readOneBC (initialize.c:400-460) + dataSplit (utilities.c:51-60) +
// dataInterp (utilities.c:65-85), extended to sample Nt
int read_bc(const frehd_case *c, const char *name, int N, std::vector<double> &out)
{
    if (N < 2) return fail(FREHD_EINVAL, "%s: the series needs at least 2 samples (size %d)", name, N);
    std::vector<double> all;
    int rc = read_values(join(c->inputFolder, name), 2 * (size_t)N, all);
    if (rc) return rc;
    std::vector<double> t(N), v(N);
    for (int q = 0; q < N; q++) {
        t[q] = all[2 * q];
        v[q] = all[2 * q + 1];
    }
    const int itval = (int)(t[1] - t[0]);
    int ind1 = 0;
    if ((rc = search_ind(t, c->info.tNStart, itval, N, &ind1))) return rc;
    const int Nt = c->cfg.Nt;
    const double dt = c->cfg.dt, t0 = c->info.tNStart;
    out.assign((size_t)Nt + 1, 0.0);
    int kk = ind1;
    for (int ii = 0; ii <= Nt; ii++) {
        const double tc = t0 + ii * dt;
        if (ii > 0 && kk + 1 < N && tc > t[kk + 1]) kk = kk + 1;
        if (kk + 1 >= N) return fail(FREHD_EINVAL, "%s ends before step %d", name, ii);
        const double slope = (v[kk + 1] - v[kk]) / (t[kk + 1] - t[kk]);
        const double intercept = v[kk] - slope * t[kk];
        out[ii] = slope * tc + intercept;
    }
    return 0;
}

int make_dirs(const std::string &path)
{
    std::string cur;
    for (size_t q = 0; q < path.size(); q++) {
        cur += path[q];
        if (path[q] == '/' && cur.size() > 1) {
            if (mkdir(cur.c_str(), 0775) != 0 && errno != EEXIST)
                return fail(FREHD_EINVAL, "cannot create %s", cur.c_str());
        }
    }
    return 0;
}

// WriteFile (fileio.c:181-206): "<folder><name>_<tt>.dat", "%4.4f \n" each
int write_file(const frehd_case *c, const char *name, int tt, const double *v, size_t N)
{
    char fn[64];
    snprintf(fn, sizeof fn, "%s_%d.dat", name, tt);
    const std::string path = c->outputFolder + fn;
    FILE *f = fopen(path.c_str(), "w");
    if (!f) return fail(FREHD_EINVAL, "cannot write %s", path.c_str());
    for (size_t q = 0; q < N; q++) fprintf(f, "%4.4f \n", v[q]);
    fclose(f);
    return 0;
}

}  // namespace

extern "C" {

// ReadInputFile (FREHD.c:28-60) + ReadUserSettings (configuration.c:18-158)
// + ReadBathymetry / InitBathymetry's offset (bathymetry.c:19-86) + readBC
// (initialize.c:348-398) + initial conditions (initialize.c:280-324) +
// initAllSubVar's tables (subgrid.c:2112-2230)
int frehd_case_load(const char *data_dat, const char *base_dir, frehd_case_t **out)
{
    if (!data_dat || !out) return fail(FREHD_EINVAL, "NULL argument");
    FILE *f = fopen(data_dat, "r");
    if (!f) return fail(FREHD_EINVAL, "cannot open %s", data_dat);
    std::vector<double> a(120, 0.0);
    std::string str[4];
    char line[150];
    int ii = 0;
    while (fgets(line, sizeof line, f) && ii < 120) {
        if (line[0] == '#') continue;
        char *tok = strtok(line, ";");
        if (!tok) tok = line;
        if (ii < 4) {
            // strings: the token up to ';' (the reference's strcpy keeps it whole)
            std::string s(tok);
            while (!s.empty() && (s.back() == '\n' || s.back() == '\r' || s.back() == ' ')) s.pop_back();
            str[ii] = s;
        } else {
            a[ii] = strtod(tok, nullptr);
        }
        ii++;
    }
    fclose(f);
    if (ii < 88) return fail(FREHD_EINVAL, "%s holds %d settings, 88 expected", data_dat, ii);

    frehd_case *c = new frehd_case();
    memset(&c->cfg, 0, sizeof c->cfg);
    memset(&c->info, 0, sizeof c->info);
    memset(&c->sub, 0, sizeof c->sub);
    const std::string base = base_dir && *base_dir ? std::string(base_dir) + "/" : std::string();
    c->inputFolder = base + str[1];
    c->outputFolder = base + str[2];
    c->subgridFolder = str[3];
    frehd_config &k = c->cfg;
    const int NX = (int)a[6], NY = (int)a[7];
    k.nx = NX;
    k.ny = NY;
    k.dx = a[4];
    k.dy = a[5];
    k.dt = a[8];
    k.Nt = (int)(a[9] * 86400 / k.dt);
    c->info.Nt = k.Nt;
    c->info.OutItvl = (int)(a[11] * 3600 / k.dt);
    c->useCellEdge = (int)a[10];
    c->info.savesurface = (int)a[12];
    c->info.saveuu = (int)a[13];
    c->info.savevv = (int)a[14];
    c->info.savedepth = (int)a[15];
    c->info.savescalar = (int)a[16];
    c->info.saveCD = (int)a[17];
    c->info.savesub = (int)a[18];
    k.bcType = (int)a[19];
    k.tideLocLengthP = (int)a[20];
    const int tideFirstP = (int)a[21], tideNP = (int)a[22];
    k.tideLocLengthM = (int)a[23];
    const int tideFirstM = (int)a[24], tideNM = (int)a[25];
    k.inflowLocLength = (int)a[26];
    const int i0 = (int)a[27], i1 = (int)a[28], j0 = (int)a[29], j1 = (int)a[30], inflowN = (int)a[31];
    k.useWind = (int)a[32];
    k.northAngle = a[33];
    const int windspdN = (int)a[34], winddirN = (int)a[35];
    k.useScalar = (int)a[36];
    k.scalarAdv = (int)a[37];
    k.inflowS = a[39];
    k.useConstTidePS = (int)a[40];
    k.tidePS = a[41];
    const int tidalPSN = (int)a[42];
    k.useConstTideMS = (int)a[43];
    k.tideMS = a[44];
    const int tidalMSN = (int)a[45];
    c->initU = a[46];
    c->initV = a[47];
    c->initSurf = a[48];
    c->initS = a[49];
    c->useConstSurf0 = (int)a[50];
    c->useConstU0 = (int)a[51];
    c->useConstV0 = (int)a[52];
    c->useConstInitS = (int)a[53];
    k.useSubgrid = (int)a[54];
    k.useSubDrag = (int)a[55];
    const double dxf = a[56];
    const double surfmax = a[58];
    k.surfmin = a[59];
    k.dsurf = a[60];
    const int staggeredV = (int)a[61];
    k.useminA = (int)a[62];
    k.beta = a[63];
    k.useminV = (int)a[64];
    c->useminV = k.useminV;
    k.g = a[67];
    k.NUx = a[68];
    k.NUy = a[69];
    k.CDnotN = (int)a[70];
    k.CDx = a[71];
    k.CDy = a[72];
    k.manningN = a[73];
    k.useThinLayer = (int)a[74];
    k.hD = a[75];
    k.CDmax = a[76];
    k.CwT = a[77];
    k.rhoa = a[78];
    k.Cw = a[79];
    k.Kx = a[80];
    k.Ky = a[81];
    k.minDepth = a[82];
    k.wtfh = a[83];
    k.CFLl = a[84];
    k.CFLh = a[85];
    k.eps = a[86];
    k.maxIter = (int)a[87] > 0 ? (int)a[87] : 1;
    k.precond = FREHD_PRECOND_JACOBI;
    k.precision = FREHD_PREC_FP64;
    k.nScalar = 1;
    k.epsInner = 1e-5;
    k.maxRefine = 8;
    c->info.useSubgrid = k.useSubgrid;
    snprintf(c->info.outputFolder, sizeof c->info.outputFolder, "%s", c->outputFolder.c_str());
    int rc = 0;
    auto bad = [&](int code, const char *msg) {
        rc = fail(code, "%s", msg);
        delete c;
        return rc;
    };
    if (NX < 1 || NY < 1) return bad(FREHD_EINVAL, "NX and NY must be positive");
    if (k.useSubgrid == 1 && staggeredV != 1) return bad(FREHD_EINVAL, "subgrid needs staggeredV = 1");
    if (k.useSubgrid == 1 && k.useSubDrag != 0 && k.useSubDrag != 2 && k.useSubDrag != 3 && k.useSubDrag != 4)
        return bad(FREHD_EINVAL, "useSubDrag must be 0, 2, 3 or 4");
    if (k.useSubgrid != 1) k.useSubDrag = 0;
    if (k.useSubgrid == 0 && c->useCellEdge == 1)
        return bad(FREHD_EINVAL, "cell-edge bathymetry on the plain grid is not supported");
    if (c->info.OutItvl < 1) return bad(FREHD_EINVAL, "output interval shorter than one step");
    if ((rc = date_num(str[0], &c->info.tNStart))) { delete c; return rc; }

    // boundary locations (configuration.c:41-61, 152-157): tideP on the top
    // row, inflow cells j*NX + i - 1 over [i0, i1) x [j0, j1)
    for (int q = 0; q < k.tideLocLengthP; q++) c->tideLocP.push_back(tideFirstP + q + k.nx * (k.ny - 1));
    for (int q = 0; q < k.tideLocLengthM; q++) c->tideLocM.push_back(tideFirstM + q);
    for (int i = i0; i < i1; i++)
        for (int j = j0; j < j1; j++) c->inflowLoc.push_back(j * NX + i - 1);
    if ((int)c->inflowLoc.size() != k.inflowLocLength)
        return bad(FREHD_EINVAL, "inflow length does not match its i/j ranges");
    for (int q : c->inflowLoc)
        if (q < 0 || q >= NX * NY) return bad(FREHD_EINVAL, "inflow cell outside the domain");

    // bathymetry (bathymetry.c:19-65) and offset (:68-79)
    const size_t N = (size_t)NX * NY;
    if (k.useSubgrid == 0) rc = read_values(join(c->inputFolder, "bath.dat"), N, c->allZ);
    else rc = read_values(join(c->inputFolder, c->subgridFolder + "bottom.dat"), N, c->allZ);
    if (!rc && k.useSubgrid != 0) {
        rc = read_values(join(c->inputFolder, c->subgridFolder + "bottomXP.dat"), N, c->allEdgeX);
        if (!rc) rc = read_values(join(c->inputFolder, c->subgridFolder + "bottomYP.dat"), N, c->allEdgeY);
    }
    if (rc) { delete c; return rc; }
    double zmin = c->allZ[0];
    for (double z : c->allZ) zmin = z < zmin ? z : zmin;
    k.offset = zmin >= 0 ? 0.0 : -zmin;
    c->info.offset = k.offset;

    // boundary series (initialize.c:348-398): inflow always read
    if (k.bcType != 2 && (rc = read_bc(c, "tideP.dat", tideNP, c->tideP))) { delete c; return rc; }
    if (k.bcType != 1 && (rc = read_bc(c, "tideM.dat", tideNM, c->tideM))) { delete c; return rc; }
    if ((rc = read_bc(c, "inflow.dat", inflowN, c->inflow))) { delete c; return rc; }
    if (k.useWind == 1 && (rc = read_bc(c, "windspd.dat", windspdN, c->windspd))) { delete c; return rc; }
    if (k.useWind == 1 && (rc = read_bc(c, "winddir.dat", winddirN, c->winddir))) { delete c; return rc; }
    if (k.useConstTidePS == 0 && (rc = read_bc(c, "salinityBC.dat", tidalPSN, c->tidalPS))) { delete c; return rc; }
    if (k.useConstTideMS == 0 && (rc = read_bc(c, "salinityBC.dat", tidalMSN, c->tidalMS))) { delete c; return rc; }

    // initial conditions (initialize.c:280-324)
    if (!c->useConstSurf0 && (rc = read_values(join(c->inputFolder, "surfic.dat"), N, c->surf0))) { delete c; return rc; }
    if (!c->useConstU0 && (rc = read_values(join(c->inputFolder, "uuic.dat"), N, c->U0))) { delete c; return rc; }
    if (!c->useConstV0 && (rc = read_values(join(c->inputFolder, "vvic.dat"), N, c->V0))) { delete c; return rc; }
    if (k.useScalar == 1 && !c->useConstInitS &&
        (rc = read_values(join(c->inputFolder, "scalaric.dat"), N, c->S0))) { delete c; return rc; }

    // subgrid tables (subgrid.c:2085-2230): Ne levels, level-major
    if (k.useSubgrid == 1) {
        k.Ne = (int)round((surfmax - k.surfmin) / k.dsurf);
        k.subR = k.dx / dxf;
        if (k.Ne < 1) return bad(FREHD_EINVAL, "subgrid level range is empty");
        const char *names[14] = {"V.dat", "Z.dat", "Vxp.dat", "Vxm.dat", "Vyp.dat", "Vym.dat", "Np.dat", "Nm.dat",
                                 "Op.dat", "Om.dat", "minVxp.dat", "minVxm.dat", "minVyp.dat", "minVym.dat"};
        const int ntab = c->useminV ? 14 : 10;
        for (int t = 0; t < ntab; t++)
            if ((rc = read_values(join(c->inputFolder, c->subgridFolder + names[t]), N * (size_t)k.Ne, c->tab[t]))) {
                delete c;
                return rc;
            }
        frehd_subgrid &s = c->sub;
        s.edgeX = c->allEdgeX.data();
        s.edgeY = c->allEdgeY.data();
        const double **dst[14] = {&s.V, &s.Z, &s.Vxp, &s.Vxm, &s.Vyp, &s.Vym, &s.Np, &s.Nm, &s.Op, &s.Om,
                                  &s.minVxp, &s.minVxm, &s.minVyp, &s.minVym};
        for (int t = 0; t < ntab; t++) *dst[t] = c->tab[t].data();
        // drag-model tables (subgrid.c:2206-2225)
        const char *dnames[3] = {"YhX.dat", "YhY.dat", "Cn.dat"};
        const double **ddst[3] = {&s.YhX, &s.YhY, &s.Cn};
        for (int t = 0; t < 3; t++) {
            const bool need = k.useSubDrag == 2 ? t < 2 : ((k.useSubDrag == 3 || k.useSubDrag == 4) && t == 2);
            if (!need) continue;
            if ((rc = read_values(join(c->inputFolder, c->subgridFolder + dnames[t]), N * (size_t)k.Ne, c->drag[t]))) {
                delete c;
                return rc;
            }
            *ddst[t] = c->drag[t].data();
        }
        // minimum / maximum face areas (subgrid.c:2160-2190)
        const char *mnames[4] = {"Nmin.dat", "Nmax.dat", "Omin.dat", "Omax.dat"};
        const double **mdst[4] = {&s.Nmin, &s.Nmax, &s.Omin, &s.Omax};
        for (int t = 0; t < 4 && k.useminA == 1; t++) {
            if ((rc = read_values(join(c->inputFolder, c->subgridFolder + mnames[t]), N * (size_t)k.Ne, c->mina[t]))) {
                delete c;
                return rc;
            }
            *mdst[t] = c->mina[t].data();
        }
    }
    *out = c;
    return 0;
}

int frehd_case_inputs(const frehd_case_t *c, frehd_config *cfg, frehd_forcing *fo, frehd_ic *ic)
{
    if (!c) return fail(FREHD_EINVAL, "case is NULL");
    if (cfg) *cfg = c->cfg;
    if (fo) {
        memset(fo, 0, sizeof *fo);
        fo->tideLocP = c->tideLocP.empty() ? nullptr : c->tideLocP.data();
        fo->tideLocM = c->tideLocM.empty() ? nullptr : c->tideLocM.data();
        fo->inflowLoc = c->inflowLoc.empty() ? nullptr : c->inflowLoc.data();
        fo->tideP = c->tideP.empty() ? nullptr : c->tideP.data();
        fo->tideM = c->tideM.empty() ? nullptr : c->tideM.data();
        fo->inflow = c->inflow.empty() ? nullptr : c->inflow.data();
        fo->tidalPS = c->tidalPS.empty() ? nullptr : c->tidalPS.data();
        fo->tidalMS = c->tidalMS.empty() ? nullptr : c->tidalMS.data();
        fo->manningCell = nullptr;
        fo->windspd = c->windspd.empty() ? nullptr : c->windspd.data();
        fo->winddir = c->winddir.empty() ? nullptr : c->winddir.data();
    }
    if (ic) {
        memset(ic, 0, sizeof *ic);
        ic->bottomZ = c->allZ.data();
        ic->surf0 = c->initSurf;
        ic->U0 = c->initU;
        ic->V0 = c->initV;
        ic->S0 = c->initS;
        ic->surf0cell = c->surf0.empty() ? nullptr : c->surf0.data();
        ic->U0cell = c->U0.empty() ? nullptr : c->U0.data();
        ic->V0cell = c->V0.empty() ? nullptr : c->V0.data();
        ic->S0cell = c->S0.empty() ? nullptr : c->S0.data();
        ic->sub = c->cfg.useSubgrid == 1 ? &c->sub : nullptr;
    }
    return 0;
}

int frehd_case_get_info(const frehd_case_t *c, frehd_case_info *info)
{
    if (!c || !info) return fail(FREHD_EINVAL, "NULL argument");
    *info = c->info;
    return 0;
}

// DataOutput (fileio.c:62-127) and SubOutput (:130-158)
int frehd_case_write_arrays(const frehd_case_t *c, int32_t tt, const double *surf, const double *uuXP,
                            const double *vvYP, const double *depth, const double *CDXP, const double *CDYP,
                            const double *S, const double *Nx, const double *Oy, const double *Vx,
                            const double *Vy, const double *V, const double *Z)
{
    if (!c) return fail(FREHD_EINVAL, "case is NULL");
    int rc = make_dirs(c->outputFolder);
    if (rc) return rc;
    const frehd_case_info &in = c->info;
    const size_t N = (size_t)c->cfg.nx * c->cfg.ny;
    if (in.saveuu == 1 && uuXP && (rc = write_file(c, "uVelocity", tt, uuXP, N))) return rc;
    if (in.savevv == 1 && vvYP && (rc = write_file(c, "vVelocity", tt, vvYP, N))) return rc;
    if (in.savesurface == 1 && surf) {
        std::vector<double> s(surf, surf + N);
        for (double &v : s) v = v - in.offset;
        if ((rc = write_file(c, "surfaceZ", tt, s.data(), N))) return rc;
    }
    if (in.savedepth == 1 && depth && (rc = write_file(c, "depth", tt, depth, N))) return rc;
    if (in.saveCD == 1 && CDXP && CDYP) {
        if ((rc = write_file(c, "cdxp", tt, CDXP, N)) || (rc = write_file(c, "cdyp", tt, CDYP, N))) return rc;
    }
    if (c->cfg.useScalar == 1 && in.savescalar == 1 && S && (rc = write_file(c, "scalar", tt, S, N))) return rc;
    if (c->cfg.useSubgrid != 0 && in.savesub == 1) {
        const char *nm[6] = {"SubNx", "SubOy", "SubVx", "SubVy", "SubV", "SubZ"};
        const double *ar[6] = {Nx, Oy, Vx, Vy, V, Z};
        for (int q = 0; q < 6; q++)
            if (ar[q] && (rc = write_file(c, nm[q], tt, ar[q], N))) return rc;
    }
    return 0;
}

int frehd_case_write_output(const frehd_case_t *c, frehd_dev_t *dev, int32_t tt)
{
    if (!c || !dev) return fail(FREHD_EINVAL, "NULL argument");
    const size_t N = (size_t)c->cfg.nx * c->cfg.ny, Ng = (size_t)(c->cfg.nx + 2) * (c->cfg.ny + 2);
    std::vector<double> surf(Ng), u(Ng), v(Ng), depth(N), cdx(Ng), cdy(Ng), S(Ng), Sm(N);
    std::vector<double> sub[6];
    frehd_fields o;
    memset(&o, 0, sizeof o);
    o.surf = surf.data();
    o.uuXP = u.data();
    o.vvYP = v.data();
    o.depth = depth.data();
    o.CDXP = cdx.data();
    o.CDYP = cdy.data();
    if (c->cfg.useScalar == 1) o.S = S.data();
    if (c->cfg.useSubgrid == 1) {
        for (auto &x : sub) x.assign(Ng, 0.0);
        o.Nx = sub[0].data();
        o.Oy = sub[1].data();
        o.Vx = sub[2].data();
        o.Vy = sub[3].data();
        o.V = sub[4].data();
        o.Z = sub[5].data();
    }
    int rc = frehd_dev_download(dev, &o);
    if (rc) return rc;
    const bool sg = c->cfg.useSubgrid == 1;
    return frehd_case_write_arrays(c, tt, o.surf, o.uuXP, o.vvYP, o.depth, o.CDXP, o.CDYP, o.S,
                                   sg ? o.Nx : nullptr, sg ? o.Oy : nullptr, sg ? o.Vx : nullptr,
                                   sg ? o.Vy : nullptr, sg ? o.V : nullptr, sg ? o.Z : nullptr);
}

int frehd_case_free(frehd_case_t *c)
{
    delete c;
    return 0;
}

}  // extern "C"
