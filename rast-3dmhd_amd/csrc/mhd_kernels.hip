// HIP kernels of the explicit compressible-MHD step for gfx950 (MI355X).
//
// K2 k_stage : right-hand side of FLUXES (3dmhdsub.f:1178-2637) fused with one
//              Wray RK3 stage update of STEP (3dmhdsub.f:926-969, 977-1067,
//              1075-1165), one launch per stage.  Stage 3 also reduces the
//              CFL limits of the new state for the next step (sub:741-873).
// K3 k_bcon  : boundary planes of BCON (3dmhdsub.f:2751-2958); after stage 3
//              it adds the boundary planes' CFL contributions.
// K1 k_cfl   : stand-alone CFL / Mach reduction for the first step after an
//              upload (sub:741-873).
// K4 k_pack / k_unpack : conversion between the reference's padded arrays
//              and the resident layout, with the periodic ghost semantics of
//              COMM_MPI (3dmhdsub.f:3131-3138).
//
// Resident layout: field f of a register is a contiguous block of (npz+2)
// planes of npx*npy doubles, x fastest.  Plane 0 and npz+1 are the top and
// bottom z ghost planes, planes 1..npz the reference's K=2..NPZ+1.  There are
// no x/y ghost cells: neighbours wrap periodically.
//
// Stage kernel organisation.  A workgroup owns a 64 x TY column tile and
// marches down a chunk of planes.  LDS holds the raw fields of planes k and
// k+1 with a one-point halo (double buffer) and a three-plane ring of the
// velocities u,v,w (needed at (i+-1,k+-1) and (j+-1,k+-1) by the cross terms
// of the viscous stress).  Plane k+2 is loaded into registers while plane k
// is computed, and the Z register of the RK3 scheme is prefetched the same
// way, so each global load has a whole plane of arithmetic to hide behind.
// Outputs are written with non-temporal stores so that L2 keeps the halo
// rows that neighbouring tiles re-read.
//
// Arithmetic: every expression keeps the reference's left-to-right
// association and the file is compiled with -ffp-contract=off, so results
// are bit-identical to the reference's FP64 build.

#include "mhd_device.h"
#include <type_traits>

namespace rmhd {

constexpr int TX = 64;  // interior columns per tile (one wavefront row)

template <int TY_>
struct Tile {
    static constexpr int TY = TY_;
    static constexpr int NT = TX * TY;      // threads per workgroup
    static constexpr int SW = TX + 2;       // LDS row stride (1-point halo each side)
    static constexpr int SH = TY + 2;
    static constexpr int SP = SW * SH;      // points per LDS plane
    static constexpr int NHALO = SP - NT;   // halo points, one per thread for t < NHALO
    static_assert(NHALO <= NT, "one halo point per thread at most");
};

// Raw-field slots of the LDS planes (resident field order).
enum { S_RU = 0, S_RV, S_RW, S_RO, S_TT, S_BX, S_BY, S_BZ };

struct Raw {
    double v[NFIELD];
};

template <bool MAG>
__device__ __forceinline__ void load_raw(const double *__restrict__ q, long long fs, long long off, Raw &r) {
#pragma unroll
    for (int f = 0; f < (MAG ? 8 : 5); ++f) r.v[f] = q[f * fs + off];
}

// CFL limits of one point (sub:750-856).  mach = OGAMMA*UU/TT before the
// square root; w1..w4 the four pointwise limits.
template <bool MAG>
__device__ __forceinline__ void cfl_point(const Params &p, const double *q, double ddx, double ddy, double ddz,
                                          double rkap, double &mach, double &w1, double &w2, double &w3,
                                          double &w4) {
    const double ru = q[F_RU], rv = q[F_RV], rw = q[F_RW], ro = q[F_RO], tt = q[F_TT];
    const double r = 1.0 / ro;
    double uu = (ru * ru + rv * rv + rw * rw) * (r * r);
    const double mm = p.ogamma * uu / tt;
    if (mm > mach || !(mm == mm)) mach = mm;  // NaN propagates into the max
    if (MAG) {
        const double bx = q[F_BX], by = q[F_BY], bz = q[F_BZ];
        const double b2 = bx * bx + by * by + bz * bz;
        uu = uu + p.gamma * tt + b2 * p.obeta / ro + 2.0 * sqrt(uu * (p.gamma * tt + b2 * p.obeta / ro));
    } else {
        uu = uu + p.gamma * tt + 2.0 * sqrt(uu * p.gamma * tt);
    }
    double dd = fmin(fmin(ddx, ddy), ddz);
    const double a1 = dd / sqrt(uu);
    const double a2 = 0.5 * dd * dd * p.repr * p.cv * ro / rkap;
    const double a3 = 0.375 * dd * dd * p.re * ro;
    w1 = (a1 < w1 || !(a1 == a1)) ? a1 : w1;
    w2 = (a2 < w2 || !(a2 == a2)) ? a2 : w2;
    w3 = (a3 < w3 || !(a3 == a3)) ? a3 : w3;
    if (MAG) w4 = fmin(w4, 0.5 * dd * dd * p.rm);
}

// Workgroup reduction of the CFL accumulators and one set of 64-bit atomics
// on order-preserving keys (min/max are exact, so the order is irrelevant).
// NaN keys sort above +Inf: they win the max and, through the key of -NaN
// (below -Inf), the min, so a non-finite state cannot go unnoticed.
__device__ void cfl_reduce(double mach, double w1, double w2, double w3, double w4, int lmag,
                           DevScalars *__restrict__ sc) {
    __shared__ unsigned long long red[16][5];
    unsigned long long km = dkey(mach), k1 = dkey(w1), k2 = dkey(w2), k3 = dkey(w3), k4 = dkey(w4);
    // a NaN limit becomes -NaN for the min reduction
    if (!(w1 == w1)) k1 = 0;
    if (!(w2 == w2)) k2 = 0;
    if (!(w3 == w3)) k3 = 0;
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o;
        o = __shfl_xor(km, off, 64); km = o > km ? o : km;
        o = __shfl_xor(k1, off, 64); k1 = o < k1 ? o : k1;
        o = __shfl_xor(k2, off, 64); k2 = o < k2 ? o : k2;
        o = __shfl_xor(k3, off, 64); k3 = o < k3 ? o : k3;
        o = __shfl_xor(k4, off, 64); k4 = o < k4 ? o : k4;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) {
        red[wv][0] = km; red[wv][1] = k1; red[wv][2] = k2; red[wv][3] = k3; red[wv][4] = k4;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < nw; ++v) {
            km = red[v][0] > km ? red[v][0] : km;
            k1 = red[v][1] < k1 ? red[v][1] : k1;
            k2 = red[v][2] < k2 ? red[v][2] : k2;
            k3 = red[v][3] < k3 ? red[v][3] : k3;
            k4 = red[v][4] < k4 ? red[v][4] : k4;
        }
        key_raise(&sc->key_mach, km);
        key_lower(&sc->key_min[0], k1);
        key_lower(&sc->key_min[1], k2);
        key_lower(&sc->key_min[2], k3);
        if (lmag) key_lower(&sc->key_min[3], k4);
    }
}

// Dynamic LDS bytes of a stage workgroup.
template <int TY, bool MAG>
constexpr int stage_lds_bytes() {
    return (2 * (MAG ? 8 : 5) + 9) * Tile<TY>::SP * (int)sizeof(double);
}

// One tile column chunk of a stage: interior columns i0..i0+TX-1, rows
// j0..j0+TY-1 (clipped to the box), planes kbeg..kend.  MODE 0 writes the
// right-hand side F to fout; MODE 1..3 performs RK3 stage s.  The whole
// workgroup calls it with the same arguments.
// OPT instantiates the optional branches of FLUXES (density diffusion,
// Rempel conduction, shear boundaries, flux relaxation, embedded heat loss;
// the per-plane tables come from mhd_opt.hip); the main path has none of
// their code.
template <int TY, bool MAG, int MODE, bool OPT = false>
__device__ __forceinline__ void stage_tile(const Params &p, const Metrics &m, const double *__restrict__ qin,
                                           double *__restrict__ qout, double *__restrict__ zr,
                                           double *__restrict__ fout, DevScalars *__restrict__ sc, int i0, int j0,
                                           int kbeg, int kend, double *lds) {
    using T = Tile<TY>;
    constexpr int SW = T::SW, SP = T::SP;
    constexpr int NF = MAG ? 8 : 5;
    double *const s_cur = lds;                  // [2][NF][SP]: raw planes k, k+1 (by parity)
    double *const s_ring = lds + 2 * NF * SP;   // [3][3][SP]: u, v, w of planes k-1, k, k+1

    const int t = threadIdx.x;
    const int npx = p.npx, npy = p.npy, npz = p.npz;
    const long long fs = p.fsize, pl = p.plane;

    // own point
    const int olx = t % TX + 1, oly = t / TX + 1;
    const int opt = oly * SW + olx;
    const int gi = i0 + olx - 1, gj = j0 + oly - 1;
    const bool valid = (gi < npx) && (gj < npy);
    const long long ooff = (long long)wrapi(gj, npy) * npx + wrapi(gi, npx);
    // halo point (t < NHALO): top row, bottom row, left column, right column
    int hpt = -1;
    long long hoff = 0;
    if (t < T::NHALO) {
        int hlx, hly;
        if (t < SW) { hlx = t; hly = 0; }
        else if (t < 2 * SW) { hlx = t - SW; hly = T::SH - 1; }
        else if (t < 2 * SW + TY) { hlx = 0; hly = t - 2 * SW + 1; }
        else { hlx = SW - 1; hly = t - 2 * SW - TY + 1; }
        hpt = hly * SW + hlx;
        hoff = (long long)wrapi(j0 + hly - 1, npy) * npx + wrapi(i0 + hlx - 1, npx);
    }
    const bool has_h = hpt >= 0;

    auto cur = [&](int k, int f, int pt) -> double & { return s_cur[((k & 1) * NF + f) * SP + pt]; };
    auto ring = [&](int k, int c, int pt) -> double & { return s_ring[((k % 3) * 3 + c) * SP + pt]; };
    auto put_cur = [&](int k, int pt, const Raw &r) {
#pragma unroll
        for (int f = 0; f < NF; ++f) cur(k, f, pt) = r.v[f];
    };
    auto put_ring = [&](int k, int pt, const Raw &r) {
        const double ri = 1.0 / r.v[S_RO];     // sub:1390
        ring(k, 0, pt) = r.v[S_RU] * ri;        // sub:1414-1424
        ring(k, 1, pt) = r.v[S_RV] * ri;
        ring(k, 2, pt) = r.v[S_RW] * ri;
    };

    Raw rm_, rp_o, rp_h;
    // warm-up: planes kbeg-1 (own values + ring), kbeg and kbeg+1 (LDS + ring)
    load_raw<MAG>(qin, fs, (long long)(kbeg - 1) * pl + ooff, rm_);
    if (OPT && p.mirror && kbeg == 1) {
        // ghost plane 0 holds RO of plane 2 while FLUXES runs (sub:1303-1304)
        rm_.v[S_RO] = qin[3 * fs + 2 * pl + ooff];
    }
    put_ring(kbeg - 1, opt, rm_);
    if (has_h) {
        load_raw<MAG>(qin, fs, (long long)(kbeg - 1) * pl + hoff, rp_h);
        if (OPT && p.mirror && kbeg == 1) rp_h.v[S_RO] = qin[3 * fs + 2 * pl + hoff];
        put_ring(kbeg - 1, hpt, rp_h);
    }
#pragma unroll 1
    for (int kk = kbeg; kk <= kbeg + 1; ++kk) {
        // ghost plane npz+1 holds RO of plane npz-1 (sub:1306-1307)
        const bool mir = OPT && p.mirror && kk == npz + 1;
        load_raw<MAG>(qin, fs, (long long)kk * pl + ooff, rp_o);
        if (mir) rp_o.v[S_RO] = qin[3 * fs + (long long)(npz - 1) * pl + ooff];
        put_cur(kk, opt, rp_o);
        put_ring(kk, opt, rp_o);
        if (has_h) {
            load_raw<MAG>(qin, fs, (long long)kk * pl + hoff, rp_h);
            if (mir) rp_h.v[S_RO] = qin[3 * fs + (long long)(npz - 1) * pl + hoff];
            put_cur(kk, hpt, rp_h);
            put_ring(kk, hpt, rp_h);
        }
    }

    const double HX = p.hx, H2X = p.h2x, HY = p.hy, H2Y = p.h2y, HZ = p.hz, H2Z = p.h2z;
    const double C43 = p.c43, OCV = p.ocv, ORE = p.ore, ORM = p.orm, OBETA = p.obeta;
    const int gix = wrapi(gi, npx) + 1, gjy = wrapi(gj, npy) + 1;
    const double dx = m.dxxdx[gix], d2x = m.d2xxdx2[gix];
    const double dy = m.dyydy[gjy], d2y = m.d2yydy2[gjy];
    const double tmpy = HY * dy;
    const double tmpy1 = HY * d2y;
    const double tmpy2 = H2Y * dy * dy;
    double coefg = 0.0, coefz = 0.0;
    if (MODE > 0) {
        const double dt = sc->dt;
        coefg = (MODE == 1 ? p.gam1 : MODE == 2 ? p.gam2 : p.gam3) * dt;
        coefz = (MODE == 1 ? p.zeta1 : p.zeta2) * dt;
    }
    // CFL accumulators of the new state (stage 3 only)
    double c_mach = 0.0, c_w1 = INFINITY, c_w2 = INFINITY, c_w3 = INFINITY, c_w4 = INFINITY;
    const double c_ddx = MODE == 3 ? m.ddx[gix] : 0.0, c_ddy = MODE == 3 ? m.ddy[gjy] : 0.0;
    // the embedded heat loss's per-column Gaussians, read once
    const double c_gx = (OPT && p.itc == 2) ? p.gx[gix] : 0.0, c_gy = (OPT && p.itc == 2) ? p.gy[gjy] : 0.0;
    // the per-column metrics arrive before the march, not inside its first plane
    settle(dx), settle(d2x), settle(dy), settle(d2y), settle(c_ddx), settle(c_ddy);
    if (OPT) settle(c_gx), settle(c_gy);

    // One plane of the march.  BND instantiates the boundary-plane form
    // (one-sided z stencils, extra planes read directly); the interior form
    // has none of that code, so its loads are waited for only where used.
    auto plane = [&](const int k, auto bnd) __attribute__((always_inline)) {
        constexpr bool BND = decltype(bnd)::value;
        const bool pre = k < kend;
        const long long o = (long long)k * pl + ooff;
        if (pre) {
            load_raw<MAG>(qin, fs, (long long)(k + 2) * pl + ooff, rp_o);
            if (has_h) load_raw<MAG>(qin, fs, (long long)(k + 2) * pl + hoff, rp_h);
        }
        // Z of this point, read once (every lane: the wrapped column is in
        // range also for lanes outside the domain)
        double zk[8];
        if (MODE >= 2) {
#pragma unroll
            for (int f = 0; f < NF; ++f) zk[f] = __builtin_nontemporal_load(&zr[f * fs + o]);
        }
        __syncthreads();
        if (valid) {
            const bool top = BND && k == 1, bot = BND && k == npz;
            const double dz = uload(m.dzzdz, k), d2z = uload(m.d2zzdz2, k);
            const double rk = uload(m.rkapa, k), dk = uload(m.dkapa, k);
            const double tmpz = HZ * dz;
            const double tmpz1 = HZ * d2z;
            const double tmpz2 = H2Z * dz * dz;
#define CUR(f, d) cur(k, f, opt + (d))
#define NXT(f) cur(k + 1, f, opt)
#define RNG(kp, c, d) ring(kp, c, opt + (d))
            double q0[8];
#pragma unroll
            for (int f = 0; f < NF; ++f) q0[f] = CUR(f, 0);
            const double ro0 = q0[S_RO], tt0 = q0[S_TT];
            const double rinv = 1.0 / ro0;
            const double u0 = RNG(k, 0, 0), v0 = RNG(k, 1, 0), w0 = RNG(k, 2, 0);
            const double um = RNG(k - 1, 0, 0), vm = RNG(k - 1, 1, 0), wm = RNG(k - 1, 2, 0);
            const double up = RNG(k + 1, 0, 0), vp = RNG(k + 1, 1, 0), wp = RNG(k + 1, 2, 0);
            const double uxp = RNG(k, 0, 1), uxm = RNG(k, 0, -1);
            const double uyp = RNG(k, 0, SW), uym = RNG(k, 0, -SW);
            const double vxp = RNG(k, 1, 1), vxm = RNG(k, 1, -1);
            const double vyp = RNG(k, 1, SW), vym = RNG(k, 1, -SW);
            const double wxp = RNG(k, 2, 1), wxm = RNG(k, 2, -1);
            const double wyp = RNG(k, 2, SW), wym = RNG(k, 2, -SW);
            // own column, planes k-1 and k+1
            const double ru_m = rm_.v[S_RU], rv_m = rm_.v[S_RV], rw_m = rm_.v[S_RW], tt_m = rm_.v[S_TT];
            const double ru_p = NXT(S_RU), rv_p = NXT(S_RV), rw_p = NXT(S_RW), tt_p = NXT(S_TT);
            const double pm = rm_.v[S_RO] * tt_m, pp = NXT(S_RO) * tt_p;
            // pressure RO*TT at the in-plane neighbours (sub:1389)
            const double pxm = CUR(S_RO, -1) * CUR(S_TT, -1), pxp = CUR(S_RO, 1) * CUR(S_TT, 1);
            const double pym = CUR(S_RO, -SW) * CUR(S_TT, -SW), pyp = CUR(S_RO, SW) * CUR(S_TT, SW);

            // extra own-column values for the one-sided forms on the boundary planes
            double rwk2 = 0.0, wk2 = 0.0, bxk2 = 0.0, bxk3 = 0.0, byk2 = 0.0, byk3 = 0.0;
            if (top || bot) {
                const int ka = top ? k + 2 : k - 2;
                const int kb = top ? k + 3 : k - 3;
                const long long oa = (long long)ka * pl + ooff, ob = (long long)kb * pl + ooff;
                rwk2 = qin[2 * fs + oa];
                wk2 = rwk2 * (1.0 / qin[3 * fs + oa]);
                if (MAG) {
                    bxk2 = qin[5 * fs + oa];
                    byk2 = qin[6 * fs + oa];
                    bxk3 = qin[5 * fs + ob];
                    byk3 = qin[6 * fs + ob];
                }
            }

            // ---- continuity (sub:1252-1295)
            double fr = (CUR(S_RU, -1) - CUR(S_RU, 1)) * HX * dx;
            fr = fr - (CUR(S_RV, SW) - CUR(S_RV, -SW)) * tmpy;
            double wz;
            if (top) wz = (4.0 * rw_p - rwk2) * tmpz;
            else if (bot) wz = (rwk2 - 4.0 * rw_m) * tmpz;
            else wz = (rw_p - rw_m) * tmpz;
            fr = fr - wz;
            if (OPT && p.id != 0) {
                // diffusion of the density perturbations W1 = RO - ROM (sub:1319-1371)
                const double romk = uload(p.rom, k);
                const double d0 = ro0 - romk;
                if (p.dh != 0.0) {
                    const double dxp = CUR(S_RO, 1) - romk, dxm = CUR(S_RO, -1) - romk;
                    const double dyp = CUR(S_RO, SW) - romk, dym = CUR(S_RO, -SW) - romk;
                    const double ww2 = (dxp - dxm) * HX * d2x + (dxp - 2.0 * d0 + dxm) * H2X * dx * dx;
                    const double ww3 = (dyp - dym) * tmpy1 + (dyp - 2.0 * d0 + dym) * tmpy2;
                    fr = fr + (ww2 + ww3) * (p.dh * ORE * uload(p.wwz, k));
                }
                if (p.dv != 0.0) {
                    const double dzp = NXT(S_RO) - uload(p.rom, k + 1), dzm = rm_.v[S_RO] - uload(p.rom, k - 1);
                    const double ww2 = (dzp - dzm) * tmpz1 + (dzp - 2.0 * d0 + dzm) * tmpz2;
                    fr = fr + ww2 * (p.dv * ORE * uload(p.wwz, k));
                }
            }

            // ---- buoyancy and pressure gradient (sub:1379, 1401-1403)
            double fw = p.grav * ro0;
            fw = fw - (pp - pm) * tmpz;
            double fv = (pym - pyp) * tmpy;
            double fu = (pxm - pxp) * HX * dx;

            // ---- momentum advection (sub:1440-1481)
            fu = fu - (CUR(S_RU, 1) * uxp - CUR(S_RU, -1) * uxm) * HX * dx;
            fu = fu - (CUR(S_RU, SW) * vyp - CUR(S_RU, -SW) * vym) * tmpy;
            fu = fu - (ru_p * wp - ru_m * wm) * tmpz;
            fv = fv - (CUR(S_RV, 1) * uxp - CUR(S_RV, -1) * uxm) * HX * dx;
            fv = fv - (CUR(S_RV, SW) * vyp - CUR(S_RV, -SW) * vym) * tmpy;
            fv = fv - (rv_p * wp - rv_m * wm) * tmpz;
            fw = fw - (CUR(S_RW, 1) * uxp - CUR(S_RW, -1) * uxm) * HX * dx;
            fw = fw - (CUR(S_RW, SW) * vyp - CUR(S_RW, -SW) * vym) * tmpy;
            fw = fw - (rw_p * wp - rw_m * wm) * tmpz;

            // ---- temperature advection (sub:1500-1505)
            const double txp = CUR(S_TT, 1), txm = CUR(S_TT, -1);
            const double typ = CUR(S_TT, SW), tym = CUR(S_TT, -SW);
            const double tzp = tt_p, tzm = tt_m;
            double ft;
            if (OPT && p.lshr) {
                // no advection; FT keeps the previous evaluation's values (sub:1493)
                ft = p.ftacc[o];
            } else {
                ft = -(u0 * (txp - txm) * HX * dx);
                ft = ft - v0 * (typ - tym) * tmpy;
                ft = ft - w0 * (tzp - tzm) * tmpz;
            }

            // ---- conduction (sub:1584-1610)
            if (OPT && p.lrem) {
                // of the perturbations W1 = TT - TTM only, plus the mean profile's
                // radiative heating HRAD (sub:1536-1581)
                const double tc = p.t_cond, tmk = uload(p.ttm, k);
                const double e0 = tt0 - tmk, exp_ = txp - tmk, exm = txm - tmk, eyp = typ - tmk, eym = tym - tmk;
                const double ezp = tzp - uload(p.ttm, k + 1), ezm = tzm - uload(p.ttm, k - 1);
                ft = ft + ((exp_ - exm) * HX * d2x + (exp_ - 2.0 * e0 + exm) * H2X * dx * dx) * rinv * tc;
                ft = ft + ((eyp - eym) * tmpy1 + (eyp - 2.0 * e0 + eym) * tmpy2) * rinv * tc;
                ft = ft + ((ezp - ezm) * tmpz1 + (ezp - 2.0 * e0 + ezm) * tmpz2) * rinv * tc;
                ft = ft + rinv * tc * uload(p.hrad, k);
            } else {
                const double tc = p.t_cond;
                ft = ft + ((txp - txm) * HX * d2x + (txp - 2.0 * tt0 + txm) * H2X * dx * dx) * rinv * tc * rk;
                ft = ft + ((typ - tym) * tmpy1 + (typ - 2.0 * tt0 + tym) * tmpy2) * rinv * tc * rk;
                ft = ft + ((tzp - tzm) * tmpz1 + (tzp - 2.0 * tt0 + tzm) * tmpz2) * rinv * tc * rk;
                ft = ft + (tzp - tzm) * HZ * dz * rinv * tc * dk;
            }
            if (OPT) {
                // flux relaxation (sub:1757-1765) and embedded heat loss (sub:1783-1806)
                if (p.rlax > 0.0) ft = ft + uload(p.radd, k);
                if (p.itc == 2) {
                    const double HH = p.hh;
                    if (p.tc != 0.0)
                        ft = ft - p.tp * OCV * 0.5 * p.ramp * rinv * c_gx / HH * c_gy / HH * uload(p.gz, k) / HH;
                    else
                        ft = ft - p.tp * OCV * rinv * c_gx / HH * c_gy / HH * uload(p.gz, k) / HH;
                }
            }
            // LSHR: no viscous or compressional heating (sub:1903-2101)
            const bool heat = !(OPT && p.lshr);

            // ---- viscous Laplacians (sub:1818-1886)
            fu = fu + ORE * (C43 * ((uxp - uxm) * HX * d2x + (uxp - 2.0 * u0 + uxm) * H2X * dx * dx)
                             + (uyp - uym) * tmpy1 + (uyp - 2.0 * u0 + uym) * tmpy2
                             + (up - um) * tmpz1 + (up - 2.0 * u0 + um) * tmpz2);
            fv = fv + ORE * ((vxp - vxm) * HX * d2x + (vxp - 2.0 * v0 + vxm) * H2X * dx * dx
                             + C43 * ((vyp - vym) * tmpy1 + (vyp - 2.0 * v0 + vym) * tmpy2)
                             + (vp - vm) * tmpz1 + (vp - 2.0 * v0 + vm) * tmpz2);
            fw = fw + ORE * ((wxp - wxm) * HX * d2x + (wxp - 2.0 * w0 + wxm) * H2X * dx * dx
                             + (wyp - wym) * tmpy1 + (wyp - 2.0 * w0 + wym) * tmpy2
                             + C43 * ((wp - wm) * tmpz1 + (wp - 2.0 * w0 + wm) * tmpz2));

            // ---- viscous and compressional heating, grad div u cross terms (sub:1891-2101)
            const double tv = p.t_visc, tv43 = p.t_visc43, tcr = p.t_cross;
            const double dxu = (uxp - uxm) * HX * dx;
            const double dxv = (vxp - vxm) * HX * dx;
            const double dxw = (wxp - wxm) * HX * dx;
            if (heat) {
                ft = ft + (C43 * dxu * dxu + dxv * dxv + dxw * dxw) * rinv * tv;
                ft = ft - OCV * dxu * tt0;
            }
            {
                const double dxv_jp = (RNG(k, 1, SW + 1) - RNG(k, 1, SW - 1)) * HX * dx;
                const double dxv_jm = (RNG(k, 1, -SW + 1) - RNG(k, 1, -SW - 1)) * HX * dx;
                fu = fu + tcr * (dxv_jp - dxv_jm) * HY * dy;
                const double dxu_jp = (RNG(k, 0, SW + 1) - RNG(k, 0, SW - 1)) * HX * dx;
                const double dxu_jm = (RNG(k, 0, -SW + 1) - RNG(k, 0, -SW - 1)) * HX * dx;
                fv = fv + tcr * (dxu_jp - dxu_jm) * HY * dy;
            }
            const double dyu = (uyp - uym) * HY * dy;
            const double dyv = (vyp - vym) * HY * dy;
            const double dzw = (wp - wm) * HZ * dz;
            if (heat) {
                ft = ft + (dyu * dyu + 2.0 * dyu * dxv) * rinv * tv;
                ft = ft + dyv * dyv * rinv * tv43;
                ft = ft - OCV * dyv * tt0;
                ft = ft + (dzw * dzw - dxu * dzw - dxu * dyv - dyv * dzw) * rinv * tv43;
                ft = ft - OCV * dzw * tt0;
            }
            {
                const double dzw_ip = (RNG(k + 1, 2, 1) - RNG(k - 1, 2, 1)) * HZ * dz;
                const double dzw_im = (RNG(k + 1, 2, -1) - RNG(k - 1, 2, -1)) * HZ * dz;
                fu = fu + tcr * (dzw_ip - dzw_im) * HX * dx;
                const double dzw_jp = (RNG(k + 1, 2, SW) - RNG(k - 1, 2, SW)) * HZ * dz;
                const double dzw_jm = (RNG(k + 1, 2, -SW) - RNG(k - 1, 2, -SW)) * HZ * dz;
                fv = fv + tcr * (dzw_jp - dzw_jm) * HY * dy;
            }
            const double dzu = (up - um) * HZ * dz;
            const double dzv = (vp - vm) * HZ * dz;
            const double dyw = (wyp - wym) * HY * dy;
            if (heat) ft = ft + (dzu * dzu + dzv * dzv + dyw * dyw + 2.0 * dzv * dyw) * rinv * tv;
            {
                const double dzu_ip = (RNG(k + 1, 0, 1) - RNG(k - 1, 0, 1)) * HZ * dz;
                const double dzu_im = (RNG(k + 1, 0, -1) - RNG(k - 1, 0, -1)) * HZ * dz;
                const double dzv_jp = (RNG(k + 1, 1, SW) - RNG(k - 1, 1, SW)) * HZ * dz;
                const double dzv_jm = (RNG(k + 1, 1, -SW) - RNG(k - 1, 1, -SW)) * HZ * dz;
                fw = fw + tcr * ((dzu_ip - dzu_im) * HX * dx + (dzv_jp - dzv_jm) * HY * dy);
            }
            if (heat) ft = ft + 2.0 * dzu * dxw * rinv * tv;

            // ---- Coriolis (sub:2106-2125)
            if (p.lrot) {
                fu = fu + p.omz * q0[S_RV];
                fv = fv - p.omz * q0[S_RU] + p.omx * q0[S_RW];
                fw = fw - p.omx * q0[S_RV];
            }

            // ---- induction, Lorentz force, Ohmic heating, resistive diffusion (sub:2129-2634)
            double w1 = 0.0, w2 = 0.0, w3 = 0.0;
            if (MAG) {
                const double bx0 = q0[S_BX], by0 = q0[S_BY], bz0 = q0[S_BZ];
                const double bxm = rm_.v[S_BX], bym = rm_.v[S_BY], bzm = rm_.v[S_BZ];
                const double bxp = NXT(S_BX), byp = NXT(S_BY), bzp = NXT(S_BZ);
                const double bxxp = CUR(S_BX, 1), bxxm = CUR(S_BX, -1);
                const double bxyp = CUR(S_BX, SW), bxym = CUR(S_BX, -SW);
                const double byxp = CUR(S_BY, 1), byxm = CUR(S_BY, -1);
                const double byyp = CUR(S_BY, SW), byym = CUR(S_BY, -SW);
                const double bzxp = CUR(S_BZ, 1), bzxm = CUR(S_BZ, -1);
                const double bzyp = CUR(S_BZ, SW), bzym = CUR(S_BZ, -SW);
                const double tm = p.t_ohm;

                w2 = bx0 * dxv - by0 * dxu;
                w3 = bx0 * dxw - bz0 * dxu;
                w1 = by0 * dyu - bx0 * dyv;
                w3 = w3 + by0 * dyw - bz0 * dyv;
                double dzw1;
                if (top) dzw1 = (4.0 * wp - wk2) * HZ * dz;
                else if (bot) dzw1 = (wk2 - 4.0 * wm) * HZ * dz;
                else dzw1 = dzw;
                w1 = w1 + bz0 * dzu - bx0 * dzw1;
                w2 = w2 + bz0 * dzv - by0 * dzw1;

                const double dxbx = (bxxp - bxxm) * HX * dx;
                const double dxby = (byxp - byxm) * HX * dx;
                const double dxbz = (bzxp - bzxm) * HX * dx;
                const double dybx = (bxyp - bxym) * HY * dy;
                fu = fu + OBETA * by0 * (dybx - dxby) - OBETA * bz0 * dxbz;
                fv = fv + OBETA * bx0 * (dxby - dybx);
                fw = fw + OBETA * bx0 * dxbz;
                ft = ft + (dybx * dybx + dxby * dxby + dxbz * dxbz - 2.0 * dxby * dybx) * rinv * tm;
                w1 = w1 - u0 * dxbx - v0 * dybx;
                w2 = w2 - u0 * dxby;
                w3 = w3 - u0 * dxbz;

                const double d2xbx = (bxxp - bxxm) * HX * d2x + (bxxp - 2.0 * bx0 + bxxm) * H2X * dx * dx;
                const double d2xby = (byxp - byxm) * HX * d2x + (byxp - 2.0 * by0 + byxm) * H2X * dx * dx;
                const double d2ybx = (bxyp - bxym) * HY * d2y + (bxyp - 2.0 * bx0 + bxym) * H2Y * dy * dy;
                w1 = w1 + ORM * (d2xbx + d2ybx);
                w2 = w2 + ORM * d2xby;
                const double d2xbz = (bzxp - bzxm) * HX * d2x + (bzxp - 2.0 * bz0 + bzxm) * H2X * dx * dx;
                w3 = w3 + ORM * d2xbz;

                const double dzbx = (bxp - bxm) * HZ * dz;
                const double dyby = (byyp - byym) * HY * dy;
                fu = fu + OBETA * bz0 * dzbx;
                fw = fw - OBETA * bx0 * dzbx;
                ft = ft + (dzbx * dzbx - 2.0 * dzbx * dxbz) * rinv * tm;
                w1 = w1 - w0 * dzbx;
                w2 = w2 - v0 * dyby;

                double d2zbx, d2zby;
                if (top) {
                    d2zbx = (-(3.0 * bx0) + 4.0 * bxp - bxk2) * HZ * d2z
                            + (2.0 * bx0 - 5.0 * bxp + 4.0 * bxk2 - bxk3) * H2Z * dz * dz;
                    d2zby = (-(3.0 * by0) + 4.0 * byp - byk2) * HZ * d2z
                            + (2.0 * by0 - 5.0 * byp + 4.0 * byk2 - byk3) * H2Z * dz * dz;
                } else if (bot) {
                    d2zbx = (3.0 * bx0 - 4.0 * bxm + bxk2) * HZ * d2z
                            + (2.0 * bx0 - 5.0 * bxm + 4.0 * bxk2 - bxk3) * H2Z * dz * dz;
                    d2zby = (3.0 * by0 - 4.0 * bym + byk2) * HZ * d2z
                            + (2.0 * by0 - 5.0 * bym + 4.0 * byk2 - byk3) * H2Z * dz * dz;
                } else {
                    d2zbx = (bxp - bxm) * HZ * d2z + (bxp - 2.0 * bx0 + bxm) * H2Z * dz * dz;
                    d2zby = (byp - bym) * HZ * d2z + (byp - 2.0 * by0 + bym) * H2Z * dz * dz;
                }
                const double d2yby = (byyp - byym) * HY * d2y + (byyp - 2.0 * by0 + byym) * H2Y * dy * dy;
                w1 = w1 + ORM * d2zbx;
                w2 = w2 + ORM * d2yby;

                const double dzby = (byp - bym) * HZ * dz;
                const double dybz = (bzyp - bzym) * HY * dy;
                const double jx = dzby - dybz;
                fv = fv + OBETA * bz0 * jx;
                fw = fw - OBETA * by0 * jx;
                ft = ft + (dzby * dzby + dybz * dybz - 2.0 * dzby * dybz) * rinv * tm;
                w2 = w2 - w0 * dzby;
                w3 = w3 - v0 * dybz;
                const double d2ybz = (bzyp - bzym) * HY * d2y + (bzyp - 2.0 * bz0 + bzym) * H2Y * dy * dy;
                w2 = w2 + ORM * d2zby;
                w3 = w3 + ORM * d2ybz;
                const double dzbz = (bzp - bzm) * HZ * dz;
                w3 = w3 - w0 * dzbz;
                const double d2zbz = (bzp - bzm) * HZ * d2z + (bzp - 2.0 * bz0 + bzm) * H2Z * dz * dz;
                w3 = w3 + ORM * d2zbz;
            }
#undef CUR
#undef NXT
#undef RNG
            // ---- output
            if (OPT && p.lshr) p.ftacc[o] = ft;
            const double F[8] = {fu, fv, fw, fr, ft, w1, w2, w3};
            if (MODE == 0) {
#pragma unroll
                for (int f = 0; f < NF; ++f) fout[f * fs + o] = F[f];
            } else {
                double qn[8];
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    // (Z is read only once this field's right-hand side exists)
                    const double zin = (MODE == 1) ? q0[f] : after(zk[f], F[f]);
                    qn[f] = zin + coefg * F[f];
                    __builtin_nontemporal_store(qn[f], &qout[f * fs + o]);
                    if (MODE < 3) __builtin_nontemporal_store(qn[f] + coefz * F[f], &zr[f * fs + o]);
                }
                if (MODE == 3 && !top && !bot)
                    cfl_point<MAG>(p, qn, c_ddx, c_ddy, uload(m.ddz, k), OPT ? uload(m.rkcfl, k) : rk, c_mach, c_w1,
                                   c_w2, c_w3, c_w4);
            }
            // own raw values of plane k become the k-1 values of the next plane
#pragma unroll
            for (int f = 0; f < NF; ++f) rm_.v[f] = q0[f];
            if (BND) settle(rwk2), settle(wk2), settle(bxk2), settle(byk2), settle(bxk3), settle(byk3);
        }
        // every load of this plane has arrived on every path out of it, so
        // none is still counted as outstanding in the next plane
        if (MODE >= 2) {
#pragma unroll
            for (int f = 0; f < NF; ++f) settle(zk[f]);
        }
        __syncthreads();
        // plane k+2 replaces plane k in LDS and enters the ring
        if (OPT && pre && p.mirror && k + 2 == npz + 1) {
            // ghost plane npz+1 holds RO of plane npz-1 = k (sub:1306-1307); each
            // thread reads the points it wrote itself
            rp_o.v[S_RO] = cur(k, S_RO, opt);
            if (has_h) rp_h.v[S_RO] = cur(k, S_RO, hpt);
        }
        if (pre) {
            put_cur(k + 2, opt, rp_o);
            put_ring(k + 2, opt, rp_o);
            if (has_h) {
                put_cur(k + 2, hpt, rp_h);
                put_ring(k + 2, hpt, rp_h);
            }
        }
    };
    int k0 = kbeg, k1 = kend;
    if (k0 == 1) plane(k0++, std::true_type{});
    const bool last_bnd = k1 == npz && k1 >= k0;
    if (last_bnd) --k1;
#pragma unroll 1
    for (int k = k0; k <= k1; ++k) plane(k, std::false_type{});
    if (last_bnd) plane(npz, std::true_type{});
    if (MODE == 3) cfl_reduce(c_mach, c_w1, c_w2, c_w3, c_w4, p.lmag, sc);
}

// Single box: workgroup -> (tile, z chunk), XCD-aware.
template <int TY, bool MAG, int MODE, bool OPT>
__global__ __launch_bounds__(TX *TY, 2) void k_stage(const Params p, const Metrics m,
                                                  const double *__restrict__ qin,
                                                  double *__restrict__ qout,
                                                  double *__restrict__ zr,
                                                  double *__restrict__ fout,
                                                  DevScalars *__restrict__ sc,
                                                  int kchunk) {
    extern __shared__ double lds[];
    const int ntx = (p.npx + TX - 1) / TX, nty = (p.npy + TY - 1) / TY;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int tz = bid / (ntx * nty);
    const int rem = bid - tz * (ntx * nty);
    const int tyb = rem / ntx, txb = rem - tyb * ntx;
    const int kbeg = 1 + tz * kchunk;
    const int kend = min(p.npz, kbeg + kchunk - 1);
    if (kbeg > p.npz) return;
    stage_tile<TY, MAG, MODE, OPT>(p, m, qin, qout, zr, fout, sc, txb * TX, tyb * TY, kbeg, kend, lds);
}

// Ensemble: a persistent grid walks the work-item table of all boxes with a
// static stride (deterministic; every workgroup leaves after its last item).
// Register a of every box is the input, 1-a the output.
template <int TY, bool MAG, int MODE>
__global__ __launch_bounds__(TX *TY, 2) void k_stage_batch(const BoxDev *__restrict__ boxes,
                                                        const StageItem *__restrict__ items, int nitems, int a) {
    extern __shared__ double lds[];
    const int slot = xcd_remap(blockIdx.x, gridDim.x);
    for (int w = slot; w < nitems; w += gridDim.x) {
        const StageItem it = items[w];
        const BoxDev &B = boxes[it.box];
        __syncthreads();  // the previous item's last LDS reads are done
        stage_tile<TY, MAG, MODE>(B.p, B.m, B.q[a], B.q[1 - a], B.z, nullptr, B.sc, it.i0, it.j0, it.kbeg, it.kend,
                                  lds);
    }
}

// BCON on the two boundary planes of the freshly written register, and the
// in-place 1/rho of the untouched z ghost planes that FLUXES performs
// (sub:1390) so that the resident ghosts follow the reference exactly.
// With CFL set (after stage 3), the boundary planes' CFL limits of the final
// state are added to the reduction.
// Column idx of the plane (npx*npy columns); the whole workgroup calls it.
template <bool MAG, bool CFL>
__device__ __forceinline__ void bcon_column(const Params &p, const Metrics &m, const double *__restrict__ qin,
                                            double *__restrict__ q, DevScalars *__restrict__ sc, int idx) {
    double c_mach = 0.0, c_w1 = INFINITY, c_w2 = INFINITY, c_w3 = INFINITY, c_w4 = INFINITY;
    if (idx < p.npx * p.npy) {
        const long long fs = p.fsize, pl = p.plane;
        const int nz = p.npz;
        const double C13 = p.c13, C23 = p.c23, C43 = p.c43;
        const int j = idx / p.npx, i = idx - j * p.npx;
        double *RU = q, *RV = q + fs, *RW = q + 2 * fs, *RO = q + 3 * fs, *TT = q + 4 * fs, *BZ = q + 7 * fs;
        // shear boundaries: RRT = 1, RRB = -1, RPP = 0 (sub:2751-2756)
        const double shear = p.lshr ? cos(0.0 * sc->timt) : 0.0;
        // top (sub:2771-2868)
        {
            const long long a = pl + idx, b = 2 * pl + idx, c = 3 * pl + idx;
            if (p.lshr) {
                RU[a] = 0.0;
                RV[a] = 1.0 * shear * RO[a];
            } else {
                RU[a] = RO[a] * (C43 * RU[b] / RO[b] - C13 * RU[c] / RO[c]);
                RV[a] = RO[a] * (C43 * RV[b] / RO[b] - C13 * RV[c] / RO[c]);
            }
            RW[a] = 0.0;
            if (p.tp != 0.0 && (p.itc == 0 || p.itc == 1)) {
                // Gaussian top temperature or flux (sub:2790-2827)
                const double HH = p.hh, ex = p.gx[i + 1], ey = p.gy[j + 1];
                if (p.itc == 0) {
                    TT[a] = 1.0 - (1.0 - p.tpr) * ex / HH * ey / HH;
                } else {
                    const double dz = p.dzf / m.dzzdz[1];
                    TT[a] = C43 * TT[b] - C13 * TT[c] - C23 * dz * (p.theta - (p.theta - p.tpr) * ex / HH * ey / HH);
                }
            } else if (p.itc == 0 || p.itc == 2 || p.itc == 4) {
                // fixed T; TU with LREM; ITC=4: TP cooler on I <= NX/2 (sub:2828-2850)
                if (p.lrem) TT[a] = p.tu;
                else TT[a] = (p.itc == 4 && i + 2 <= (p.npx + 2) / 2) ? 1.0 - p.tp : 1.0;
            } else if (p.itc == 1) {
                if (p.lrem) {
                    TT[a] = C43 * TT[b] - C13 * TT[c] + p.dztu;
                } else {
                    const double dz = p.dzf / m.dzzdz[1];
                    TT[a] = C43 * TT[b] - C13 * TT[c] - C23 * dz * p.theta;
                }
            }  // ITC=3: the top temperature is left to the equations
            if (MAG && p.ibc == 0) BZ[a] = 0.0;
        }
        // bottom (sub:2873-2955)
        {
            const long long a = (long long)nz * pl + idx, b = (long long)(nz - 1) * pl + idx,
                            c = (long long)(nz - 2) * pl + idx;
            if (p.lshr) {
                RU[a] = 0.0;
                RV[a] = -1.0 * shear * RO[a];
            } else {
                RU[a] = RO[a] * (C43 * RU[b] / RO[b] - C13 * RU[c] / RO[c]);
                RV[a] = RO[a] * (C43 * RV[b] / RO[b] - C13 * RV[c] / RO[c]);
            }
            RW[a] = 0.0;
            if (p.izc == 0) {
                TT[a] = p.tb;
            } else if (p.lrem) {
                TT[a] = C43 * TT[b] - C13 * TT[c] + p.dztb;
            } else {
                const double dz = p.dzf / m.dzzdz[nz];
                TT[a] = C43 * TT[b] - C13 * TT[c] + C23 * dz * p.theta / m.rkapa[nz];
            }
            if (MAG && p.ibc == 0) BZ[a] = 0.0;
        }
        // z ghost planes of RO flip to 1/RO once per FLUXES call; with ID or
        // LREM they held the mirrored RO of planes 2 and npz-1 (sub:1301-1310)
        const double *ROi = qin + 3 * fs;
        RO[idx] = 1.0 / ROi[(p.mirror ? 2 * pl : 0) + idx];
        RO[(long long)(nz + 1) * pl + idx] = 1.0 / ROi[(long long)(p.mirror ? nz - 1 : nz + 1) * pl + idx];
        if (CFL) {
            for (int kb = 0; kb < 2; ++kb) {
                const int k = kb ? nz : 1;
                const long long o = (long long)k * pl + idx;
                double v[8];
#pragma unroll
                for (int f = 0; f < (MAG ? 8 : 5); ++f) v[f] = q[f * fs + o];
                cfl_point<MAG>(p, v, m.ddx[i + 1], m.ddy[j + 1], m.ddz[k], m.rkcfl[k], c_mach, c_w1, c_w2,
                               c_w3, c_w4);
            }
        }
    }
    if (CFL) cfl_reduce(c_mach, c_w1, c_w2, c_w3, c_w4, p.lmag, sc);
}

template <bool MAG, bool CFL>
__global__ __launch_bounds__(256) void k_bcon(const Params p, const Metrics m, const double *__restrict__ qin,
                                              double *__restrict__ q, DevScalars *__restrict__ sc) {
    bcon_column<MAG, CFL>(p, m, qin, q, sc, blockIdx.x * blockDim.x + threadIdx.x);
}

// Ensemble: one workgroup per (box, run of 256 columns); register a of
// every box was the stage input, 1-a the freshly written output.
template <bool MAG, bool CFL>
__global__ __launch_bounds__(256) void k_bcon_batch(const BoxDev *__restrict__ boxes,
                                                    const ColumnItem *__restrict__ items, int a) {
    const ColumnItem it = items[blockIdx.x];
    const BoxDev &B = boxes[it.box];
    bcon_column<MAG, CFL>(B.p, B.m, B.q[a], B.q[1 - a], B.sc, it.col0 + threadIdx.x);
}

// Stand-alone CFL reduction over planes 1..npz (first step after an upload).
template <bool MAG>
__global__ __launch_bounds__(256) void k_cfl(const Params p, const Metrics m, const double *__restrict__ q,
                                             DevScalars *__restrict__ sc) {
    const long long fs = p.fsize, pl = p.plane;
    const long long n = (long long)p.npz * pl;
    double mach = 0.0, w1 = INFINITY, w2 = INFINITY, w3 = INFINITY, w4 = INFINITY;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long o = pl + idx;
        const int k = (int)(idx / pl) + 1;
        const long long r2 = idx - (long long)(k - 1) * pl;
        const int j = (int)(r2 / p.npx), i = (int)(r2 - (long long)j * p.npx);
        double v[8];
#pragma unroll
        for (int f = 0; f < (MAG ? 8 : 5); ++f) v[f] = q[f * fs + o];
        cfl_point<MAG>(p, v, m.ddx[i + 1], m.ddy[j + 1], m.ddz[k], m.rkcfl[k], mach, w1, w2, w3, w4);
    }
    cfl_reduce(mach, w1, w2, w3, w4, p.lmag, sc);
}

// End of a reduction: DT = SF*MIN(...), UMACH = SQRT(MAX(...)) (sub:761,
// 868-872), then re-arm the keys.  A positive fixed_dt replaces the CFL DT.  With clock set this also closes the step
// that just ran: NIT, TIMC, TIMT (sub:1171-1173) with the DT it used, which
// is kept as dt_used / umach_used.
__device__ __forceinline__ void cfl_final(DevScalars *sc, double sf, int lmag, int clock) {
    if (clock) {
        sc->nit = sc->nit + 1;
        sc->timc = sc->timc + sc->dt;
        sc->timt = sc->timi + sc->timc;
        sc->dt_used = sc->dt;
        sc->umach_used = sc->umach;
    }
    const double mach = dunkey(sc->key_mach);
    double mn = INFINITY;
    bool bad = !(mach == mach);
    for (int n = 0; n < 4; ++n) {
        sc->cflmin[n] = dunkey(sc->key_min[n]);
        if (n < 3 || lmag) {
            mn = fmin(mn, sc->cflmin[n]);
            bad = bad || !(sc->cflmin[n] == sc->cflmin[n]);
        }
    }
    sc->umach = sqrt(mach);
    const double dcfl = sf * mn;
    sc->dt = sc->fixed_dt > 0.0 ? sc->fixed_dt : dcfl;
    if (bad || dcfl == INFINITY) sc->nonfinite |= 1;
    sc->key_mach = dkey(0.0);
    for (int n = 0; n < 4; ++n) sc->key_min[n] = dkey(INFINITY);
}

__global__ void k_cfl_final(DevScalars *sc, double sf, int lmag, int clock) { cfl_final(sc, sf, lmag, clock); }

// Ensemble: one thread per box.
__global__ void k_cfl_final_batch(const BoxDev *__restrict__ boxes, int nbox, int clock) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbox) cfl_final(boxes[b].sc, boxes[b].p.sf, boxes[b].p.lmag, clock);
}

// padded Fortran (npx+2, npy+2, npz+2) -> resident field (x/y ghosts dropped)
__global__ void k_pack(const Params p, const double *__restrict__ padded, double *__restrict__ field) {
    const long long n = p.plane * (p.npz + 2);  // the field's values, not its padded stride
    const int nxp = p.npx + 2, nyp = p.npy + 2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long k = idx / p.plane;
        const long long r = idx - k * p.plane;
        const int j = (int)(r / p.npx), i = (int)(r - (long long)j * p.npx);
        field[idx] = padded[(k * nyp + (j + 1)) * nxp + (i + 1)];
    }
}

// resident field -> padded Fortran array; periodic ghost fill (mode 1) or
// zero ghosts and zero z ghost planes (mode 0, for right-hand sides)
__global__ void k_unpack(const Params p, const double *__restrict__ field, double *__restrict__ padded, int mode) {
    const int nxp = p.npx + 2, nyp = p.npy + 2, nzp = p.npz + 2;
    const long long n = (long long)nxp * nyp * nzp;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long k = idx / ((long long)nxp * nyp);
        const long long r = idx - k * nxp * nyp;
        const int J = (int)(r / nxp), I = (int)(r - (long long)J * nxp);
        if (mode == 0) {
            const bool inner = I >= 1 && I <= p.npx && J >= 1 && J <= p.npy && k >= 1 && k <= p.npz;
            padded[idx] = inner ? field[k * p.plane + (long long)(J - 1) * p.npx + (I - 1)] : 0.0;
        } else {
            const int i = wrapi(I - 1, p.npx), j = wrapi(J - 1, p.npy);
            padded[idx] = field[k * p.plane + (long long)j * p.npx + i];
        }
    }
}

// Box filter of nf resident fields (planes 1..npz) onto a coarse grid:
// out[f][K][J][I], I fastest.  Horizontally the mean of factor x factor
// blocks (coarse I covers fine i = factor*I .. factor*I + factor-1).
// Vertically, zmode 0: the mean of factor planes (coarse K covers fine
// planes factor*K+1 .. factor*K+factor); zmode 1 (vertex grids with npz =
// factor*(nzc-1)+1): coarse K sits on fine plane factor*K+1 and averages the
// planes within factor/2 of it with trapezoid weights (1/2 at both ends);
// the two boundary planes are taken as they are.
__global__ void k_box_filter(const Params p, const double *__restrict__ q, int nf, int factor, int zmode, int nxc,
                             int nyc, int nzc, double *__restrict__ out) {
    const long long n = (long long)nf * nzc * nyc * nxc;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int I = (int)(e % nxc);
        long long r = e / nxc;
        const int J = (int)(r % nyc);
        r /= nyc;
        const int K = (int)(r % nzc);
        const int f = (int)(r / nzc);
        const double *qf = q + (long long)f * p.fsize;
        auto hmean = [&](int k) {
            double s = 0.0;
            for (int jj = 0; jj < factor; ++jj) {
                const double *row = qf + (long long)k * p.plane + (long long)(J * factor + jj) * p.npx + I * factor;
                for (int ii = 0; ii < factor; ++ii) s += row[ii];
            }
            return s / (double)(factor * factor);
        };
        double v = 0.0;
        if (zmode == 0) {
            for (int kk = 1; kk <= factor; ++kk) v += hmean(K * factor + kk);
            v /= (double)factor;
        } else if (K == 0 || K == nzc - 1) {
            v = hmean(K * factor + 1);
        } else {
            const int c = K * factor + 1, hw = factor / 2;
            for (int d = -hw; d <= hw; ++d) {
                const double w = (d == -hw || d == hw) ? 0.5 : 1.0;
                v += w * hmean(c + d);
            }
            v /= (double)factor;
        }
        out[e] = v;
    }
}

// ---------------------------------------------------------------------------
// launchers

// z chunks of about 32 planes: the tiles a cohort of co-resident workgroups
// (one XCD) marches stay closer in step over a short march, so more of the
// halo rows they share are still in L2 when the neighbour reads them
// (512x512x256: 8 chunks fetch 14% fewer bytes and run 3% faster than 2
// chunks of 128 planes).  At least 2048 workgroups; at least 8 planes.
static int chunk_for(const Params &p, int ty, int *nchunks) {
    const int ntiles = ((p.npx + TX - 1) / TX) * ((p.npy + ty - 1) / ty);
    int nc = (p.npz + 31) / 32;
    const int need = (2048 + ntiles - 1) / ntiles;     // workgroups to fill the chip
    if (nc < need) nc = need;
    int maxc = p.npz / 8;                              // but keep >= 8 planes per chunk
    if (maxc < 1) maxc = 1;
    if (nc > maxc) nc = maxc;
    if (nc < 1) nc = 1;
    const int kc = (p.npz + nc - 1) / nc;
    *nchunks = (p.npz + kc - 1) / kc;
    return kc;
}

template <int TY, bool MAG, int MODE, bool OPT>
static hipError_t launch_stage_o(const Params &p, const Metrics &m, const double *qin, double *qout, double *z,
                                 double *fout, DevScalars *sc, hipStream_t s) {
    constexpr int lds = stage_lds_bytes<TY, MAG>();
    static std::atomic<unsigned long long> attr{0};
    {
        hipError_t e = lds_limit_once(attr, reinterpret_cast<const void *>(&k_stage<TY, MAG, MODE, OPT>), lds);
        if (e != hipSuccess) return e;
    }
    int nch;
    const int kc = chunk_for(p, TY, &nch);
    const int ntiles = ((p.npx + TX - 1) / TX) * ((p.npy + TY - 1) / TY);
    hipLaunchKernelGGL((k_stage<TY, MAG, MODE, OPT>), dim3(ntiles * nch), dim3(TX * TY), lds, s, p, m, qin, qout,
                       z, fout, sc, kc);
    return hipGetLastError();
}

template <int TY, bool MAG, int MODE>
static hipError_t launch_stage_t(const Params &p, const Metrics &m, const double *qin, double *qout, double *z,
                                 double *fout, DevScalars *sc, hipStream_t s) {
    if (p.opt) return launch_stage_o<TY, MAG, MODE, true>(p, m, qin, qout, z, fout, sc, s);
    return launch_stage_o<TY, MAG, MODE, false>(p, m, qin, qout, z, fout, sc, s);
}

template <int TY>
static hipError_t launch_stage_ty(int mode, const Params &p, const Metrics &m, const double *qin, double *qout,
                                  double *z, double *fout, DevScalars *sc, hipStream_t s) {
    if (p.lmag) {
        switch (mode) {
            case 0: return launch_stage_t<TY, true, 0>(p, m, qin, qout, z, fout, sc, s);
            case 1: return launch_stage_t<TY, true, 1>(p, m, qin, qout, z, fout, sc, s);
            case 2: return launch_stage_t<TY, true, 2>(p, m, qin, qout, z, fout, sc, s);
            default: return launch_stage_t<TY, true, 3>(p, m, qin, qout, z, fout, sc, s);
        }
    }
    switch (mode) {
        case 0: return launch_stage_t<TY, false, 0>(p, m, qin, qout, z, fout, sc, s);
        case 1: return launch_stage_t<TY, false, 1>(p, m, qin, qout, z, fout, sc, s);
        case 2: return launch_stage_t<TY, false, 2>(p, m, qin, qout, z, fout, sc, s);
        default: return launch_stage_t<TY, false, 3>(p, m, qin, qout, z, fout, sc, s);
    }
}

hipError_t launch_stage(int mode, const Params &p, const Metrics &m, const double *qin, double *qout,
                        double *z, double *fout, DevScalars *sc, hipStream_t s) {
    if (p.tile_y == 8) return launch_stage_ty<8>(mode, p, m, qin, qout, z, fout, sc, s);
    return launch_stage_ty<4>(mode, p, m, qin, qout, z, fout, sc, s);
}

template <int TY, bool MAG, int MODE>
static hipError_t launch_batch_t(const BoxDev *boxes, const StageItem *items, int nitems, int grid, int a,
                                 hipStream_t s) {
    constexpr int lds = stage_lds_bytes<TY, MAG>();
    static std::atomic<unsigned long long> attr{0};
    {
        hipError_t e = lds_limit_once(attr, reinterpret_cast<const void *>(&k_stage_batch<TY, MAG, MODE>), lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_stage_batch<TY, MAG, MODE>), dim3(grid), dim3(TX * TY), lds, s, boxes, items, nitems, a);
    return hipGetLastError();
}

template <int TY>
static hipError_t launch_batch_ty(int mode, int lmag, const BoxDev *boxes, const StageItem *items, int nitems,
                                  int grid, int a, hipStream_t s) {
    if (lmag) {
        if (mode == 1) return launch_batch_t<TY, true, 1>(boxes, items, nitems, grid, a, s);
        if (mode == 2) return launch_batch_t<TY, true, 2>(boxes, items, nitems, grid, a, s);
        return launch_batch_t<TY, true, 3>(boxes, items, nitems, grid, a, s);
    }
    if (mode == 1) return launch_batch_t<TY, false, 1>(boxes, items, nitems, grid, a, s);
    if (mode == 2) return launch_batch_t<TY, false, 2>(boxes, items, nitems, grid, a, s);
    return launch_batch_t<TY, false, 3>(boxes, items, nitems, grid, a, s);
}

hipError_t launch_stage_batch(int mode, int tile_y, int lmag, const BoxDev *boxes, const StageItem *items,
                              int nitems, int grid, int a, hipStream_t s) {
    if (tile_y == 8) return launch_batch_ty<8>(mode, lmag, boxes, items, nitems, grid, a, s);
    return launch_batch_ty<4>(mode, lmag, boxes, items, nitems, grid, a, s);
}

hipError_t launch_bcon_batch(int lmag, const BoxDev *boxes, const ColumnItem *items, int nitems, int a, bool cfl,
                             hipStream_t s) {
    dim3 grid(nitems), block(256);
    if (lmag) {
        if (cfl) hipLaunchKernelGGL((k_bcon_batch<true, true>), grid, block, 0, s, boxes, items, a);
        else hipLaunchKernelGGL((k_bcon_batch<true, false>), grid, block, 0, s, boxes, items, a);
    } else {
        if (cfl) hipLaunchKernelGGL((k_bcon_batch<false, true>), grid, block, 0, s, boxes, items, a);
        else hipLaunchKernelGGL((k_bcon_batch<false, false>), grid, block, 0, s, boxes, items, a);
    }
    return hipGetLastError();
}

hipError_t launch_cfl_final_batch(const BoxDev *boxes, int nbox, bool clock, hipStream_t s) {
    hipLaunchKernelGGL(k_cfl_final_batch, dim3((nbox + 63) / 64), dim3(64), 0, s, boxes, nbox, clock ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_bcon(const Params &p, const Metrics &m, const double *qin, double *qout, DevScalars *sc,
                       bool cfl, hipStream_t s) {
    const int n = p.npx * p.npy;
    dim3 grid((n + 255) / 256), block(256);
    if (p.lmag) {
        if (cfl) hipLaunchKernelGGL((k_bcon<true, true>), grid, block, 0, s, p, m, qin, qout, sc);
        else hipLaunchKernelGGL((k_bcon<true, false>), grid, block, 0, s, p, m, qin, qout, sc);
    } else {
        if (cfl) hipLaunchKernelGGL((k_bcon<false, true>), grid, block, 0, s, p, m, qin, qout, sc);
        else hipLaunchKernelGGL((k_bcon<false, false>), grid, block, 0, s, p, m, qin, qout, sc);
    }
    return hipGetLastError();
}

hipError_t launch_cfl(const Params &p, const Metrics &m, const double *q, DevScalars *sc, hipStream_t s) {
    const long long n = (long long)p.npz * p.plane;
    long long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    dim3 grid((unsigned)nb), block(256);
    if (p.lmag) hipLaunchKernelGGL(k_cfl<true>, grid, block, 0, s, p, m, q, sc);
    else hipLaunchKernelGGL(k_cfl<false>, grid, block, 0, s, p, m, q, sc);
    return hipGetLastError();
}

hipError_t launch_cfl_final(const Params &p, DevScalars *sc, bool clock, hipStream_t s) {
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(1), 0, s, sc, p.sf, p.lmag, clock ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_pack(const Params &p, const double *padded, double *field, hipStream_t s) {
    long long nb = (p.plane * (p.npz + 2) + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)nb), dim3(256), 0, s, p, padded, field);
    return hipGetLastError();
}

hipError_t launch_unpack(const Params &p, const double *field, double *padded, int mode, hipStream_t s) {
    const long long n = (long long)(p.npx + 2) * (p.npy + 2) * (p.npz + 2);
    long long nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_unpack, dim3((unsigned)nb), dim3(256), 0, s, p, field, padded, mode);
    return hipGetLastError();
}

hipError_t launch_box_filter(const Params &p, const double *q, int nf, int factor, int zmode, int nxc, int nyc, int nzc,
                             double *out, hipStream_t s) {
    const long long n = (long long)nf * nzc * nyc * nxc;
    long long nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_box_filter, dim3((unsigned)nb), dim3(256), 0, s, p, q, nf, factor, zmode, nxc, nyc, nzc, out);
    return hipGetLastError();
}

}  // namespace rmhd
