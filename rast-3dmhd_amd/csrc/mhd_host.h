// Host-side helpers shared by the single-box and the ensemble C-ABI.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/rast3dmhd.h"
#include "mhd_kernels.h"

namespace rmhd {

// Records an error for mhd_last_error(NULL) (the message of calls without a
// handle) and returns code (mhd_capi.hip).
int global_fail(int code, const char *fmt, ...);

// Whether a descriptor switches on any optional branch of FLUXES/BCON.
inline bool desc_branches(const mhd_desc *d) {
    return d->id != 0 || d->lrem || d->lshr || d->rlax > 0.0 || d->itcon > 1 || d->tp != 0.0;
}

// Branches whose kernels read the coordinate tables (EXX, WYY, ZEE).
inline bool desc_needs_coords(const mhd_desc *d) {
    return (d->tp != 0.0 && (d->itcon == 0 || d->itcon == 1)) || d->itcon == 2 || d->id == 2 || d->rlax > 0.0;
}

// Checks a descriptor; on failure writes a message to msg and returns the
// MHD_ERR_* code.
inline int check_desc(const mhd_desc *d, char *msg, size_t len) {
    if (d->npx < 3 || d->npy < 3 || d->npz < 4) {
        snprintf(msg, len, "box %dx%dx%d too small (need npx,npy>=3, npz>=4)", d->npx, d->npy, d->npz);
        return MHD_ERR_ARG;
    }
    if (d->izcon != 0 && d->izcon != 1) {
        snprintf(msg, len, "IZCON=%d (BCON: Invalid IZCON)", d->izcon);
        return MHD_ERR_SCOPE;
    }
    if (d->itcon < 0 || d->itcon > 4) {
        snprintf(msg, len, "ITCON=%d not supported (0..4)", d->itcon);
        return MHD_ERR_SCOPE;
    }
    if (d->precision != MHD_FP64 && d->precision != MHD_MIXED) {
        snprintf(msg, len, "precision %d unknown", d->precision);
        return MHD_ERR_ARG;
    }
    if (d->id < 0 || d->id > 2) {
        snprintf(msg, len, "ID=%d not supported (0..2)", d->id);
        return MHD_ERR_SCOPE;
    }
    if (d->rlax > 0.0 && !(d->itcon == 0 && d->izcon == 1)) {
        // the reference stops for other pairs (sub:1642-1646), and its fixed-flux
        // top / fixed-T bottom integral reads FRADM(NZ), which it never sets on
        // the bottom rank (sub:1679-1691, 1745-1748)
        snprintf(msg, len, "flux relaxation (RLAX > 0) needs ITCON=0 and IZCON=1");
        return MHD_ERR_SCOPE;
    }
    if (d->lshr && (d->lrem || d->rlax > 0.0)) {
        snprintf(msg, len, "LSHR with LREM or RLAX (the reference uses a stale TTM there)");
        return MHD_ERR_SCOPE;
    }
    const int npey = d->npey > 0 ? d->npey : 1, npez = d->npez > 0 ? d->npez : 1;
    if (d->npy % npey || d->npz % npez) {
        snprintf(msg, len, "NPEY=%d / NPEZ=%d do not divide NPY=%d / NPZ=%d", npey, npez, d->npy, d->npz);
        return MHD_ERR_ARG;
    }
    if (desc_branches(d) && d->precision != MHD_FP64) {
        snprintf(msg, len, "the optional FLUXES/BCON branches (ID, LREM, LSHR, RLAX, ITCON 2..4, TP) are built "
                           "in the MHD_FP64 form only");
        return MHD_ERR_SCOPE;
    }
    return MHD_OK;
}

// Kernel parameters of a box.  The products the reference forms once per
// FLUXES / STEP call are formed here with the same association.
inline Params params_from_desc(const mhd_desc *d) {
    Params p{};
    p.npx = d->npx; p.npy = d->npy; p.npz = d->npz;
    p.lmag = d->lmag ? 1 : 0; p.lrot = d->lrot ? 1 : 0;
    p.izc = d->izcon; p.itc = d->itcon; p.ibc = d->ibcon;
    p.tile_y = 4;
    if (const char *ty = getenv("RAST3DMHD_TILE_Y")) p.tile_y = atoi(ty) == 8 ? 8 : 4;
    p.plane = (long long)d->npx * d->npy;
    p.fsize = p.plane * (d->npz + 2) + FIELD_PAD;
    p.hx = d->hx; p.h2x = d->h2x; p.hy = d->hy; p.h2y = d->h2y; p.hz = d->hz; p.h2z = d->h2z;
    p.c13 = d->c13; p.c23 = d->c23; p.c43 = d->c43;
    p.gam1 = d->gam1; p.gam2 = d->gam2; p.gam3 = d->gam3; p.zeta1 = d->zeta1; p.zeta2 = d->zeta2;
    p.cv = d->cv; p.ocv = d->ocv; p.ore = d->ore; p.re = d->re; p.repr = d->repr;
    p.theta = d->theta; p.grav = d->grav; p.sf = d->sf; p.gamma = d->gamma;
    p.omx = d->lrot ? d->omx : 0.0; p.omz = d->lrot ? d->omz : 0.0;
    p.orm = d->orm; p.rm = d->rm; p.obeta = d->obeta; p.tb = d->tb; p.dzf = d->dzf;
    p.ogamma = 1.0 / d->gamma;             // sub:744
    p.t_cond = d->ocv / d->repr;           // sub:1584
    p.t_visc = d->ocv * d->ore;            // sub:1904
    p.t_visc43 = d->c43 * d->ocv * d->ore; // sub:1962
    p.t_cross = d->c13 * d->ore;           // sub:1930
    p.t_ohm = d->obeta * d->orm * d->ocv;  // sub:2287
    p.id = d->id; p.lrem = d->lrem ? 1 : 0; p.lshr = d->lshr ? 1 : 0;
    p.mirror = (d->id != 0 || d->lrem) ? 1 : 0;
    p.npey = d->npey > 0 ? d->npey : 1; p.npez = d->npez > 0 ? d->npez : 1;
    p.dh = d->dh; p.dv = d->dv; p.tu = d->tu; p.dztu = d->dztu; p.dztb = d->dztb;
    p.rlax = d->rlax; p.tstart = d->tstart; p.toff = d->toff; p.tp = d->tp; p.tc = d->tc; p.hh = d->hh;
    p.tpr = d->tp;
    p.opt = (p.id != 0 || p.lrem || p.lshr || p.rlax > 0.0 || p.itc == 2) ? 1 : 0;
    return p;
}

// Initial device scalars: reduction keys armed, clocks zero.
inline DevScalars initial_scalars() {
    DevScalars s{};
    s.key_mach = dkey(0.0);
    for (int n = 0; n < 4; ++n) s.key_min[n] = dkey(INFINITY);
    s.dd2min = INFINITY;
    return s;
}

// min over the points i = 1..npx, j = 1..npy, k = 1..npz of
// min(DDX_i, DDY_j, DDZ_k)^2, the grid-only term of the magnetic CFL limit
// (the same squares and minimum the per-point reduction forms).
inline double dd2_min(const Params &p, const double *ddx, const double *ddy, const double *ddz) {
    double m = INFINITY;
    for (int i = 1; i <= p.npx; ++i) m = std::fmin(m, ddx[i] * ddx[i]);
    for (int j = 1; j <= p.npy; ++j) m = std::fmin(m, ddy[j] * ddy[j]);
    for (int k = 1; k <= p.npz; ++k) m = std::fmin(m, ddz[k] * ddz[k]);
    return m;
}

// Number of doubles of the 11 metric arrays of a box.
inline size_t metric_len(const Params &p) {
    return 3 * (size_t)(p.npx + 2) + 3 * (size_t)(p.npy + 2) + 5 * (size_t)(p.npz + 2);
}

// Lengths of the 11 metric arrays in the order dxxdx, d2xxdx2, ddx, dyydy,
// d2yydy2, ddy, dzzdz, d2zzdz2, ddz, rkapa, dkapa.
inline void metric_lengths(const Params &p, size_t len[11]) {
    const size_t lx = p.npx + 2, ly = p.npy + 2, lz = p.npz + 2;
    const size_t l[11] = {lx, lx, lx, ly, ly, ly, lz, lz, lz, lz, lz};
    for (int n = 0; n < 11; ++n) len[n] = l[n];
}

// Points the Metrics table into a packed device block of metric_len doubles
// (same order) and returns each array's length in len[11].
inline Metrics metric_table(const Params &p, const double *base, size_t len[11]) {
    metric_lengths(p, len);
    const double *a[11];
    size_t off = 0;
    for (int n = 0; n < 11; ++n) {
        a[n] = base + off;
        off += len[n];
    }
    Metrics m;
    m.dxxdx = a[0]; m.d2xxdx2 = a[1]; m.ddx = a[2];
    m.dyydy = a[3]; m.d2yydy2 = a[4]; m.ddy = a[5];
    m.dzzdz = a[6]; m.d2zzdz2 = a[7]; m.ddz = a[8];
    m.rkapa = a[9]; m.dkapa = a[10];
    m.rkcfl = a[9];
    return m;
}

// Host tables of the mixed form's background for planes 0..npz+1:
// dbl = [ro | tt | fwbg | ftbg], flt = [ro32 | tt32 | dt0dz32 | rinv0_32 | s0_32].
// The background actually used is the given profile rounded to FP32
// values, so that the FP32 copies the perturbation stencil multiplies with
// are exact: the perturbation pressure rho0 T' + rho' (T0 + T') is a small
// residual of nearly cancelling terms in low-Mach flow, and different
// rounding errors of rho0 and T0 would leave a systematic error in it.
// fwbg, ftbg are the background's own discrete residuals of the momentum z
// and temperature equations (conduction S0 / rho0), formed in FP64 with the
// same stencils as the step.  iddz2 = 1/ddz^2 and irk = 1/rkapa are per-plane
// factors of the CFL limits, divided here once instead of per point (IEEE
// division on host and device alike).  Returns the first plane with
// rho0 <= 0, or -1.
constexpr int BG_DBL = 6, BG_FLT = 5;  // FP64 and FP32 tables of npz+2 entries
inline int background_tables(const Params &p, const double *dzzdz, const double *d2zzdz2, const double *ddz,
                             const double *rkapa, const double *dkapa, const double *ro0, const double *tt0,
                             std::vector<double> &dbl, std::vector<float> &flt, std::vector<PlaneM> &pl) {
    const int nz = p.npz + 2;
    dbl.assign(BG_DBL * (size_t)nz, 0.0);
    flt.assign(BG_FLT * (size_t)nz, 0.0f);
    double *ro = dbl.data(), *tt = ro + nz, *fw = tt + nz, *ft = fw + nz, *iz = ft + nz, *irk = iz + nz;
    float *ro32 = flt.data(), *tt32 = ro32 + nz, *dtz = tt32 + nz, *ri = dtz + nz, *s0 = ri + nz;
    for (int k = 0; k < nz; ++k) {
        if (!(ro0[k] > 0.0) || !((float)ro0[k] > 0.0f)) return k;
        ro32[k] = (float)ro0[k];
        tt32[k] = (float)tt0[k];
        ro[k] = (double)ro32[k];
        tt[k] = (double)tt32[k];
        ri[k] = (float)(1.0 / ro[k]);
    }
    for (int k = 1; k <= p.npz; ++k) {
        const double dz = dzzdz[k], d2z = d2zzdz2[k];
        const double g0 = (tt[k + 1] - tt[k - 1]) * p.hz * dz;
        const double l0 = (tt[k + 1] - tt[k - 1]) * p.hz * d2z + (tt[k + 1] - 2.0 * tt[k] + tt[k - 1]) * p.h2z * dz * dz;
        const double sz = p.t_cond * (rkapa[k] * l0 + dkapa[k] * g0);
        dtz[k] = (float)g0;
        s0[k] = (float)sz;
        ft[k] = sz / ro[k];
        fw[k] = p.grav * ro[k] - (ro[k + 1] * tt[k + 1] - ro[k - 1] * tt[k - 1]) * p.hz * dz;
        iz[k] = 1.0 / (ddz[k] * ddz[k]);
        irk[k] = 1.0 / rkapa[k];
    }
    pl.assign((size_t)nz, PlaneM{});
    for (int k = 0; k < nz; ++k) {
        PlaneM &q = pl[k];
        q.dzzdz = dzzdz[k];
        q.d2zzdz2 = d2zzdz2[k];
        q.rkapa = rkapa[k];
        q.dkapa = dkapa[k];
        q.ddz = ddz[k];
        q.ro = ro[k];
        q.tt = tt[k];
        q.fwbg = fw[k];
        q.ftbg = ft[k];
        q.iddz2 = iz[k];
        q.irk = irk[k];
        q.ro32 = ro32[k];
        q.tt32 = tt32[k];
        q.dt0dz32 = dtz[k];
        q.rinv0_32 = ri[k];
        q.s0_32 = s0[k];
        q.c43 = (float)p.c43;
        q.ocv = (float)p.ocv;
        q.ore = (float)p.ore;
        q.orm = (float)p.orm;
        q.obeta = (float)p.obeta;
        q.grav = (float)p.grav;
        q.t_cond = (float)p.t_cond;
        q.t_visc = (float)p.t_visc;
        q.t_visc43 = (float)p.t_visc43;
        q.t_cross = (float)p.t_cross;
        q.t_ohm = (float)p.t_ohm;
        q.omx = (float)p.omx;
        q.omz = (float)p.omz;
    }
    return -1;
}

// Background table pointers into device blocks laid out as background_tables.
inline Background background_view(int npz, const double *bgd, const float *bgf) {
    const int nz = npz + 2;
    Background bg{};
    bg.ro = bgd;
    bg.tt = bgd + nz;
    bg.fwbg = bgd + 2 * nz;
    bg.ftbg = bgd + 3 * nz;
    bg.iddz2 = bgd + 4 * nz;
    bg.irk = bgd + 5 * nz;
    bg.ro32 = bgf;
    bg.tt32 = bgf + nz;
    bg.dt0dz32 = bgf + 2 * nz;
    bg.rinv0_32 = bgf + 3 * nz;
    bg.s0_32 = bgf + 4 * nz;
    return bg;
}

}  // namespace rmhd
