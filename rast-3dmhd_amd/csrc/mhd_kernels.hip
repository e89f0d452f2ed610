// HIP kernels of the explicit compressible-MHD step for gfx950 (MI355X).
//
// K2 k_stage : right-hand side of FLUXES (3dmhdsub.f:1178-2637) fused with one
//              Wray RK3 stage update of STEP (3dmhdsub.f:926-969, 977-1067,
//              1075-1165), one launch per stage.
// K3 k_bcon  : boundary planes of BCON (3dmhdsub.f:2751-2958).
// K1 k_cfl   : pointwise CFL limits and maximum Mach number with exact
//              min/max reductions (3dmhdsub.f:741-873).
// K4 k_pack / k_unpack : conversion between the reference's padded arrays
//              and the resident layout, with the periodic ghost semantics of
//              COMM_MPI (3dmhdsub.f:3131-3138).
//
// Resident layout: field f of a register is a contiguous block of
// (npz+2) planes of npx*npy doubles, x fastest.  Plane 0 and npz+1 are the
// top/bottom z ghost planes, planes 1..npz the reference's K=2..NPZ+1.
// There are no x/y ghost cells: neighbours wrap periodically.
//
// Arithmetic: every expression keeps the reference's left-to-right
// association and the file is compiled with -ffp-contract=off, so results
// are bit-identical to the reference's FP64 build.

#include "mhd_kernels.h"

namespace rmhd {

constexpr int TX = 64;              // interior columns per tile (one wavefront row)
constexpr int TY = 4;               // interior rows per tile
constexpr int NT = TX * TY;         // threads per workgroup
constexpr int SW = TX + 2;          // LDS row stride (1-point halo each side)
constexpr int SH = TY + 2;
constexpr int SP = SW * SH;         // points per LDS plane
constexpr int NHALO = SP - NT;      // halo points, one per thread for t < NHALO
static_assert(NHALO <= NT, "one halo point per thread at most");

// LDS slots of the current plane.
enum { C_RU = 0, C_RV, C_RW, C_TT, C_P, C_BX, C_BY, C_BZ };

struct Raw {
    double ru, rv, rw, ro, tt, bx, by, bz;
};

template <bool MAG>
__device__ __forceinline__ void load_raw(const double *__restrict__ q, long long fs, long long off, Raw &r) {
    r.ru = q[off];
    r.rv = q[fs + off];
    r.rw = q[2 * fs + off];
    r.ro = q[3 * fs + off];
    r.tt = q[4 * fs + off];
    if (MAG) {
        r.bx = q[5 * fs + off];
        r.by = q[6 * fs + off];
        r.bz = q[7 * fs + off];
    } else {
        r.bx = r.by = r.bz = 0.0;
    }
}

__device__ __forceinline__ int wrapi(int x, int n) {
    int r = x % n;
    return r < 0 ? r + n : r;
}

// XCD-aware block order: the hardware deals consecutive workgroup ids to the
// 8 XCDs in turn; give each XCD a contiguous run of tiles so that tiles that
// share halo rows share an L2.
__device__ __forceinline__ int xcd_remap(int bid, int nb) {
    if ((nb & 7) != 0) return bid;
    return (bid & 7) * (nb >> 3) + (bid >> 3);
}

// MODE 0: write the right-hand side F to fout.  MODE 1..3: RK3 stage.
template <bool MAG, int MODE>
__global__ __launch_bounds__(NT) void k_stage(const Params p, const Metrics m,
                                              const double *__restrict__ qin,
                                              double *__restrict__ qout,
                                              double *__restrict__ zr,
                                              double *__restrict__ fout,
                                              const DevScalars *__restrict__ sc,
                                              int kchunk) {
    constexpr int NC = MAG ? 8 : 5;
    __shared__ double s_cur[NC][SP];     // rho u, rho v, rho w, T, P (, B) of plane k
    __shared__ double s_ring[3][3][SP];  // u, v, w of planes k-1, k, k+1 (slot = plane % 3)

    const int t = threadIdx.x;
    const int npx = p.npx, npy = p.npy, npz = p.npz;
    const long long fs = p.fsize, pl = p.plane;
    const int ntx = (npx + TX - 1) / TX, nty = (npy + TY - 1) / TY;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int tz = bid / (ntx * nty);
    const int rem = bid - tz * (ntx * nty);
    const int tyb = rem / ntx, txb = rem - tyb * ntx;
    const int i0 = txb * TX, j0 = tyb * TY;
    const int kbeg = 1 + tz * kchunk;
    const int kend = min(npz, kbeg + kchunk - 1);
    if (kbeg > npz) return;

    // own point
    const int olx = t % TX + 1, oly = t / TX + 1;
    const int opt = oly * SW + olx;
    const int gi = i0 + olx - 1, gj = j0 + oly - 1;
    const bool valid = (gi < npx) && (gj < npy);
    const long long ooff = (long long)wrapi(gj, npy) * npx + wrapi(gi, npx);
    // halo point (t < NHALO): top row, bottom row, left column, right column
    int hpt = -1;
    long long hoff = 0;
    if (t < NHALO) {
        int hlx, hly;
        if (t < SW) { hlx = t; hly = 0; }
        else if (t < 2 * SW) { hlx = t - SW; hly = SH - 1; }
        else if (t < 2 * SW + TY) { hlx = 0; hly = t - 2 * SW + 1; }
        else { hlx = SW - 1; hly = t - 2 * SW - TY + 1; }
        hpt = hly * SW + hlx;
        hoff = (long long)wrapi(j0 + hly - 1, npy) * npx + wrapi(i0 + hlx - 1, npx);
    }
    const bool has_h = hpt >= 0;

    auto put_cur = [&](int pt, const Raw &r) {
        s_cur[C_RU][pt] = r.ru;
        s_cur[C_RV][pt] = r.rv;
        s_cur[C_RW][pt] = r.rw;
        s_cur[C_TT][pt] = r.tt;
        s_cur[C_P][pt] = r.ro * r.tt;           // pressure (sub:1389)
        if (MAG) {
            s_cur[C_BX][pt] = r.bx;
            s_cur[C_BY][pt] = r.by;
            s_cur[C_BZ][pt] = r.bz;
        }
    };
    auto put_ring = [&](int slot, int pt, const Raw &r) {
        const double ri = 1.0 / r.ro;          // sub:1390
        s_ring[slot][0][pt] = r.ru * ri;        // sub:1414-1424
        s_ring[slot][1][pt] = r.rv * ri;
        s_ring[slot][2][pt] = r.rw * ri;
    };

    Raw rm_, rc, rn_o, rn_h, rp_o, rp_h;
    // warm-up: planes kbeg-1, kbeg, kbeg+1
    {
        Raw a;
        load_raw<MAG>(qin, fs, (long long)(kbeg - 1) * pl + ooff, rm_);
        put_ring((kbeg - 1) % 3, opt, rm_);
        if (has_h) {
            load_raw<MAG>(qin, fs, (long long)(kbeg - 1) * pl + hoff, a);
            put_ring((kbeg - 1) % 3, hpt, a);
        }
        load_raw<MAG>(qin, fs, (long long)kbeg * pl + ooff, rc);
        put_cur(opt, rc);
        put_ring(kbeg % 3, opt, rc);
        if (has_h) {
            load_raw<MAG>(qin, fs, (long long)kbeg * pl + hoff, a);
            put_cur(hpt, a);
            put_ring(kbeg % 3, hpt, a);
        }
        load_raw<MAG>(qin, fs, (long long)(kbeg + 1) * pl + ooff, rn_o);
        put_ring((kbeg + 1) % 3, opt, rn_o);
        if (has_h) {
            load_raw<MAG>(qin, fs, (long long)(kbeg + 1) * pl + hoff, rn_h);
            put_ring((kbeg + 1) % 3, hpt, rn_h);
        }
    }

    const double HX = p.hx, H2X = p.h2x, HY = p.hy, H2Y = p.h2y, HZ = p.hz, H2Z = p.h2z;
    const double C43 = p.c43, OCV = p.ocv, ORE = p.ore, ORM = p.orm, OBETA = p.obeta;
    // per-column metrics
    const double dx = m.dxxdx[wrapi(gi, npx) + 1], d2x = m.d2xxdx2[wrapi(gi, npx) + 1];
    const double dy = m.dyydy[wrapi(gj, npy) + 1], d2y = m.d2yydy2[wrapi(gj, npy) + 1];
    const double tmpy = HY * dy;
    const double tmpy1 = HY * d2y;
    const double tmpy2 = H2Y * dy * dy;
    double dt = 0.0, coefg = 0.0, coefz = 0.0;
    if (MODE > 0) {
        dt = sc->dt;
        coefg = (MODE == 1 ? p.gam1 : MODE == 2 ? p.gam2 : p.gam3) * dt;
        coefz = (MODE == 1 ? p.zeta1 : p.zeta2) * dt;
    }

    for (int k = kbeg; k <= kend; ++k) {
        const bool pre = k < kend;
        if (pre) {
            load_raw<MAG>(qin, fs, (long long)(k + 2) * pl + ooff, rp_o);
            if (has_h) load_raw<MAG>(qin, fs, (long long)(k + 2) * pl + hoff, rp_h);
        }
        __syncthreads();
        if (valid) {
            const int sm = (k + 2) % 3, s0 = k % 3, sp = (k + 1) % 3;
            const bool top = (k == 1), bot = (k == npz);
            const double dz = m.dzzdz[k], d2z = m.d2zzdz2[k];
            const double rk = m.rkapa[k], dk = m.dkapa[k];
            const double tmpz = HZ * dz;
            const double tmpz1 = HZ * d2z;
            const double tmpz2 = H2Z * dz * dz;
#define CUR(f, d) s_cur[f][opt + (d)]
#define RNG(s, c, d) s_ring[s][c][opt + (d)]
            const double ro0 = rc.ro, tt0 = rc.tt;
            const double rinv = 1.0 / ro0;
            const double u0 = RNG(s0, 0, 0), v0 = RNG(s0, 1, 0), w0 = RNG(s0, 2, 0);
            const double um = RNG(sm, 0, 0), vm = RNG(sm, 1, 0), wm = RNG(sm, 2, 0);
            const double up = RNG(sp, 0, 0), vp = RNG(sp, 1, 0), wp = RNG(sp, 2, 0);
            const double uxp = RNG(s0, 0, 1), uxm = RNG(s0, 0, -1);
            const double uyp = RNG(s0, 0, SW), uym = RNG(s0, 0, -SW);
            const double vxp = RNG(s0, 1, 1), vxm = RNG(s0, 1, -1);
            const double vyp = RNG(s0, 1, SW), vym = RNG(s0, 1, -SW);
            const double wxp = RNG(s0, 2, 1), wxm = RNG(s0, 2, -1);
            const double wyp = RNG(s0, 2, SW), wym = RNG(s0, 2, -SW);
            const double pm = rm_.ro * rm_.tt, pp = rn_o.ro * rn_o.tt;

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
            double fr = (CUR(C_RU, -1) - CUR(C_RU, 1)) * HX * dx;
            fr = fr - (CUR(C_RV, SW) - CUR(C_RV, -SW)) * tmpy;
            double wz;
            if (top) wz = (4.0 * rn_o.rw - rwk2) * tmpz;
            else if (bot) wz = (rwk2 - 4.0 * rm_.rw) * tmpz;
            else wz = (rn_o.rw - rm_.rw) * tmpz;
            fr = fr - wz;

            // ---- buoyancy and pressure gradient (sub:1379, 1401-1403)
            double fw = p.grav * ro0;
            fw = fw - (pp - pm) * tmpz;
            double fv = (CUR(C_P, -SW) - CUR(C_P, SW)) * tmpy;
            double fu = (CUR(C_P, -1) - CUR(C_P, 1)) * HX * dx;

            // ---- momentum advection (sub:1440-1481)
            fu = fu - (CUR(C_RU, 1) * uxp - CUR(C_RU, -1) * uxm) * HX * dx;
            fu = fu - (CUR(C_RU, SW) * vyp - CUR(C_RU, -SW) * vym) * tmpy;
            fu = fu - (rn_o.ru * wp - rm_.ru * wm) * tmpz;
            fv = fv - (CUR(C_RV, 1) * uxp - CUR(C_RV, -1) * uxm) * HX * dx;
            fv = fv - (CUR(C_RV, SW) * vyp - CUR(C_RV, -SW) * vym) * tmpy;
            fv = fv - (rn_o.rv * wp - rm_.rv * wm) * tmpz;
            fw = fw - (CUR(C_RW, 1) * uxp - CUR(C_RW, -1) * uxm) * HX * dx;
            fw = fw - (CUR(C_RW, SW) * vyp - CUR(C_RW, -SW) * vym) * tmpy;
            fw = fw - (rn_o.rw * wp - rm_.rw * wm) * tmpz;

            // ---- temperature advection (sub:1500-1505)
            const double txp = CUR(C_TT, 1), txm = CUR(C_TT, -1);
            const double typ = CUR(C_TT, SW), tym = CUR(C_TT, -SW);
            const double tzp = rn_o.tt, tzm = rm_.tt;
            double ft = -(u0 * (txp - txm) * HX * dx);
            ft = ft - v0 * (typ - tym) * tmpy;
            ft = ft - w0 * (tzp - tzm) * tmpz;

            // ---- conduction (sub:1584-1610)
            {
                const double tc = p.t_cond;
                ft = ft + ((txp - txm) * HX * d2x + (txp - 2.0 * tt0 + txm) * H2X * dx * dx) * rinv * tc * rk;
                ft = ft + ((typ - tym) * tmpy1 + (typ - 2.0 * tt0 + tym) * tmpy2) * rinv * tc * rk;
                ft = ft + ((tzp - tzm) * tmpz1 + (tzp - 2.0 * tt0 + tzm) * tmpz2) * rinv * tc * rk;
                ft = ft + (tzp - tzm) * HZ * dz * rinv * tc * dk;
            }

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
            ft = ft + (C43 * dxu * dxu + dxv * dxv + dxw * dxw) * rinv * tv;
            ft = ft - OCV * dxu * tt0;
            {
                const double dxv_jp = (RNG(s0, 1, SW + 1) - RNG(s0, 1, SW - 1)) * HX * dx;
                const double dxv_jm = (RNG(s0, 1, -SW + 1) - RNG(s0, 1, -SW - 1)) * HX * dx;
                fu = fu + tcr * (dxv_jp - dxv_jm) * HY * dy;
                const double dxu_jp = (RNG(s0, 0, SW + 1) - RNG(s0, 0, SW - 1)) * HX * dx;
                const double dxu_jm = (RNG(s0, 0, -SW + 1) - RNG(s0, 0, -SW - 1)) * HX * dx;
                fv = fv + tcr * (dxu_jp - dxu_jm) * HY * dy;
            }
            const double dyu = (uyp - uym) * HY * dy;
            ft = ft + (dyu * dyu + 2.0 * dyu * dxv) * rinv * tv;
            const double dyv = (vyp - vym) * HY * dy;
            ft = ft + dyv * dyv * rinv * tv43;
            ft = ft - OCV * dyv * tt0;
            const double dzw = (wp - wm) * HZ * dz;
            ft = ft + (dzw * dzw - dxu * dzw - dxu * dyv - dyv * dzw) * rinv * tv43;
            ft = ft - OCV * dzw * tt0;
            {
                const double dzw_ip = (RNG(sp, 2, 1) - RNG(sm, 2, 1)) * HZ * dz;
                const double dzw_im = (RNG(sp, 2, -1) - RNG(sm, 2, -1)) * HZ * dz;
                fu = fu + tcr * (dzw_ip - dzw_im) * HX * dx;
                const double dzw_jp = (RNG(sp, 2, SW) - RNG(sm, 2, SW)) * HZ * dz;
                const double dzw_jm = (RNG(sp, 2, -SW) - RNG(sm, 2, -SW)) * HZ * dz;
                fv = fv + tcr * (dzw_jp - dzw_jm) * HY * dy;
            }
            const double dzu = (up - um) * HZ * dz;
            const double dzv = (vp - vm) * HZ * dz;
            const double dyw = (wyp - wym) * HY * dy;
            ft = ft + (dzu * dzu + dzv * dzv + dyw * dyw + 2.0 * dzv * dyw) * rinv * tv;
            {
                const double dzu_ip = (RNG(sp, 0, 1) - RNG(sm, 0, 1)) * HZ * dz;
                const double dzu_im = (RNG(sp, 0, -1) - RNG(sm, 0, -1)) * HZ * dz;
                const double dzv_jp = (RNG(sp, 1, SW) - RNG(sm, 1, SW)) * HZ * dz;
                const double dzv_jm = (RNG(sp, 1, -SW) - RNG(sm, 1, -SW)) * HZ * dz;
                fw = fw + tcr * ((dzu_ip - dzu_im) * HX * dx + (dzv_jp - dzv_jm) * HY * dy);
            }
            ft = ft + 2.0 * dzu * dxw * rinv * tv;

            // ---- Coriolis (sub:2106-2125)
            if (p.lrot) {
                fu = fu + p.omz * rc.rv;
                fv = fv - p.omz * rc.ru + p.omx * rc.rw;
                fw = fw - p.omx * rc.rv;
            }

            // ---- induction, Lorentz force, Ohmic heating, resistive diffusion (sub:2129-2634)
            double w1 = 0.0, w2 = 0.0, w3 = 0.0;
            if (MAG) {
                const double bx0 = rc.bx, by0 = rc.by, bz0 = rc.bz;
                const double bxm = rm_.bx, bym = rm_.by, bzm = rm_.bz;
                const double bxp = rn_o.bx, byp = rn_o.by, bzp = rn_o.bz;
                const double bxxp = CUR(C_BX, 1), bxxm = CUR(C_BX, -1);
                const double bxyp = CUR(C_BX, SW), bxym = CUR(C_BX, -SW);
                const double byxp = CUR(C_BY, 1), byxm = CUR(C_BY, -1);
                const double byyp = CUR(C_BY, SW), byym = CUR(C_BY, -SW);
                const double bzxp = CUR(C_BZ, 1), bzxm = CUR(C_BZ, -1);
                const double bzyp = CUR(C_BZ, SW), bzym = CUR(C_BZ, -SW);
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
                const double cur = dzby - dybz;
                fv = fv + OBETA * bz0 * cur;
                fw = fw - OBETA * by0 * cur;
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
#undef RNG
            // ---- output
            const long long o = (long long)k * pl + ooff;
            const double fv8[8] = {fu, fv, fw, fr, ft, w1, w2, w3};
            if (MODE == 0) {
#pragma unroll
                for (int f = 0; f < (MAG ? 8 : 5); ++f) fout[f * fs + o] = fv8[f];
            } else {
                const double q8[8] = {rc.ru, rc.rv, rc.rw, rc.ro, rc.tt, rc.bx, rc.by, rc.bz};
#pragma unroll
                for (int f = 0; f < (MAG ? 8 : 5); ++f) {
                    const double zin = (MODE == 1) ? q8[f] : zr[f * fs + o];
                    const double qn = zin + coefg * fv8[f];
                    qout[f * fs + o] = qn;
                    if (MODE < 3) zr[f * fs + o] = qn + coefz * fv8[f];
                }
            }
        }
        __syncthreads();
        // advance the window: k+1 becomes current, k+2 enters the ring
        rm_ = rc;
        rc = rn_o;
        put_cur(opt, rn_o);
        if (has_h) put_cur(hpt, rn_h);
        if (pre) {
            put_ring((k + 2) % 3, opt, rp_o);
            if (has_h) put_ring((k + 2) % 3, hpt, rp_h);
            rn_o = rp_o;
            rn_h = rp_h;
        }
    }
}

// BCON on the two boundary planes of the freshly written register, and the
// in-place 1/rho of the untouched z ghost planes that FLUXES performs
// (sub:1390) so that the resident ghosts follow the reference exactly.
template <bool MAG>
__global__ __launch_bounds__(256) void k_bcon(const Params p, const Metrics m, const double *__restrict__ qin,
                                              double *__restrict__ q) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= p.npx * p.npy) return;
    const long long fs = p.fsize, pl = p.plane;
    const int nz = p.npz;
    const double C13 = p.c13, C23 = p.c23, C43 = p.c43;
    double *RU = q, *RV = q + fs, *RW = q + 2 * fs, *RO = q + 3 * fs, *TT = q + 4 * fs, *BZ = q + 7 * fs;
    // top (sub:2771-2868)
    {
        const long long a = pl + idx, b = 2 * pl + idx, c = 3 * pl + idx;
        RU[a] = RO[a] * (C43 * RU[b] / RO[b] - C13 * RU[c] / RO[c]);
        RV[a] = RO[a] * (C43 * RV[b] / RO[b] - C13 * RV[c] / RO[c]);
        RW[a] = 0.0;
        if (p.itc == 0) {
            TT[a] = 1.0;
        } else {
            const double dz = p.dzf / m.dzzdz[1];
            TT[a] = C43 * TT[b] - C13 * TT[c] - C23 * dz * p.theta;
        }
        if (MAG && p.ibc == 0) BZ[a] = 0.0;
    }
    // bottom (sub:2873-2955)
    {
        const long long a = (long long)nz * pl + idx, b = (long long)(nz - 1) * pl + idx,
                        c = (long long)(nz - 2) * pl + idx;
        RU[a] = RO[a] * (C43 * RU[b] / RO[b] - C13 * RU[c] / RO[c]);
        RV[a] = RO[a] * (C43 * RV[b] / RO[b] - C13 * RV[c] / RO[c]);
        RW[a] = 0.0;
        if (p.izc == 0) {
            TT[a] = p.tb;
        } else {
            const double dz = p.dzf / m.dzzdz[nz];
            TT[a] = C43 * TT[b] - C13 * TT[c] + C23 * dz * p.theta / m.rkapa[nz];
        }
        if (MAG && p.ibc == 0) BZ[a] = 0.0;
    }
    // z ghost planes of RO flip to 1/RO once per FLUXES call
    const double *ROi = qin + 3 * fs;
    RO[idx] = 1.0 / ROi[idx];
    RO[(long long)(nz + 1) * pl + idx] = 1.0 / ROi[(long long)(nz + 1) * pl + idx];
}

// wave64 + workgroup reductions of the CFL limits (sub:747-872)
template <bool MAG>
__global__ __launch_bounds__(256) void k_cfl(const Params p, const Metrics m, const double *__restrict__ q,
                                             DevScalars *__restrict__ sc) {
    const long long fs = p.fsize, pl = p.plane;
    const long long n = (long long)p.npz * pl;
    double mach = 0.0, w1 = INFINITY, w2 = INFINITY, w3 = INFINITY, w4 = INFINITY;
    int bad = 0;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long o = pl + idx;      // planes 1..npz
        const int k = (int)(idx / pl) + 1;
        const long long r2 = idx - (long long)(k - 1) * pl;
        const int j = (int)(r2 / p.npx), i = (int)(r2 - (long long)j * p.npx);
        const double ru = q[o], rv = q[fs + o], rw = q[2 * fs + o], ro = q[3 * fs + o], tt = q[4 * fs + o];
        const double r = 1.0 / ro;
        double uu = (ru * ru + rv * rv + rw * rw) * (r * r);
        const double mm = p.ogamma * uu / tt;
        if (mm > mach) mach = mm;
        if (MAG) {
            const double bx = q[5 * fs + o], by = q[6 * fs + o], bz = q[7 * fs + o];
            const double b2 = bx * bx + by * by + bz * bz;
            uu = uu + p.gamma * tt + b2 * p.obeta / ro + 2.0 * sqrt(uu * (p.gamma * tt + b2 * p.obeta / ro));
        } else {
            uu = uu + p.gamma * tt + 2.0 * sqrt(uu * p.gamma * tt);
        }
        double dd = m.ddx[i + 1];
        dd = fmin(dd, m.ddy[j + 1]);
        dd = fmin(dd, m.ddz[k]);
        const double a1 = dd / sqrt(uu);
        const double a2 = 0.5 * dd * dd * p.repr * p.cv * ro / m.rkapa[k];
        const double a3 = 0.375 * dd * dd * p.re * ro;
        if (!(a1 == a1) || !(a2 == a2) || !(a3 == a3) || !(mm == mm)) bad = 1;
        w1 = fmin(w1, a1);
        w2 = fmin(w2, a2);
        w3 = fmin(w3, a3);
        if (MAG) w4 = fmin(w4, 0.5 * dd * dd * p.rm);
    }
    // wave64 butterfly
    for (int off = 32; off > 0; off >>= 1) {
        mach = fmax(mach, __shfl_xor(mach, off, 64));
        w1 = fmin(w1, __shfl_xor(w1, off, 64));
        w2 = fmin(w2, __shfl_xor(w2, off, 64));
        w3 = fmin(w3, __shfl_xor(w3, off, 64));
        w4 = fmin(w4, __shfl_xor(w4, off, 64));
        bad |= __shfl_xor(bad, off, 64);
    }
    __shared__ double red[4][5];
    __shared__ int rbad[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[wv][0] = mach; red[wv][1] = w1; red[wv][2] = w2; red[wv][3] = w3; red[wv][4] = w4;
        rbad[wv] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < (int)(blockDim.x >> 6); ++v) {
            mach = fmax(mach, red[v][0]);
            w1 = fmin(w1, red[v][1]);
            w2 = fmin(w2, red[v][2]);
            w3 = fmin(w3, red[v][3]);
            w4 = fmin(w4, red[v][4]);
            bad |= rbad[v];
        }
        atomicMax(&sc->key_mach, dkey(mach));
        atomicMin(&sc->key_min[0], dkey(w1));
        atomicMin(&sc->key_min[1], dkey(w2));
        atomicMin(&sc->key_min[2], dkey(w3));
        if (MAG) atomicMin(&sc->key_min[3], dkey(w4));
        if (bad) atomicOr(&sc->nonfinite, 1);
    }
}

// DT = SF*MIN(...), UMACH = SQRT(MAX(...)) (sub:761, 868-872); re-arms the keys.
__global__ void k_cfl_final(DevScalars *sc, double sf, int lmag) {
    const double mach = dunkey(sc->key_mach);
    double mn = INFINITY;
    for (int n = 0; n < 4; ++n) {
        sc->cflmin[n] = dunkey(sc->key_min[n]);
        if (n < 3 || lmag) mn = fmin(mn, sc->cflmin[n]);
    }
    sc->umach = sqrt(mach);
    sc->dt = sf * mn;
    if (!(sc->dt == sc->dt) || !(sc->umach == sc->umach) || sc->dt == INFINITY) sc->nonfinite |= 2;
    sc->key_mach = dkey(0.0);
    for (int n = 0; n < 4; ++n) sc->key_min[n] = dkey(INFINITY);
}

// NIT, TIMC, TIMT bookkeeping at the end of STEP (sub:1171-1173).
__global__ void k_clock(DevScalars *sc) {
    sc->nit = sc->nit + 1;
    sc->timc = sc->timc + sc->dt;
    sc->timt = sc->timi + sc->timc;
}

// padded Fortran (npx+2, npy+2, npz+2) -> resident field (x/y ghosts dropped)
__global__ void k_pack(const Params p, const double *__restrict__ padded, double *__restrict__ field) {
    const long long n = p.fsize;
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

// ---------------------------------------------------------------------------
// launchers

static int chunk_for(const Params &p, int *nchunks) {
    const int ntiles = ((p.npx + TX - 1) / TX) * ((p.npy + TY - 1) / TY);
    int nc = (2048 + ntiles - 1) / ntiles;            // aim for >= 2048 workgroups
    int maxc = p.npz / 8;                              // but keep >= 8 planes per chunk
    if (maxc < 1) maxc = 1;
    if (nc > maxc) nc = maxc;
    if (nc < 1) nc = 1;
    const int kc = (p.npz + nc - 1) / nc;
    *nchunks = (p.npz + kc - 1) / kc;
    return kc;
}

hipError_t launch_stage(int mode, const Params &p, const Metrics &m, const double *qin, double *qout,
                        double *z, double *fout, DevScalars *sc, hipStream_t s) {
    int nch;
    const int kc = chunk_for(p, &nch);
    const int ntiles = ((p.npx + TX - 1) / TX) * ((p.npy + TY - 1) / TY);
    dim3 grid(ntiles * nch), block(NT);
#define L(MAG, MODE) hipLaunchKernelGGL((k_stage<MAG, MODE>), grid, block, 0, s, p, m, qin, qout, z, fout, sc, kc)
    if (p.lmag) {
        switch (mode) {
            case 0: L(true, 0); break;
            case 1: L(true, 1); break;
            case 2: L(true, 2); break;
            default: L(true, 3); break;
        }
    } else {
        switch (mode) {
            case 0: L(false, 0); break;
            case 1: L(false, 1); break;
            case 2: L(false, 2); break;
            default: L(false, 3); break;
        }
    }
#undef L
    return hipGetLastError();
}

hipError_t launch_bcon(const Params &p, const Metrics &m, const double *qin, double *qout, hipStream_t s) {
    const int n = p.npx * p.npy;
    dim3 grid((n + 255) / 256), block(256);
    if (p.lmag) hipLaunchKernelGGL(k_bcon<true>, grid, block, 0, s, p, m, qin, qout);
    else hipLaunchKernelGGL(k_bcon<false>, grid, block, 0, s, p, m, qin, qout);
    return hipGetLastError();
}

hipError_t launch_cfl(const Params &p, const Metrics &m, const double *q, DevScalars *sc, hipStream_t s) {
    const long long n = (long long)p.npz * p.plane;
    long long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    dim3 grid((unsigned)nb), block(256);
    if (p.lmag) hipLaunchKernelGGL(k_cfl<true>, grid, block, 0, s, p, m, q, sc);
    else hipLaunchKernelGGL(k_cfl<false>, grid, block, 0, s, p, m, q, sc);
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(1), 0, s, sc, p.sf, p.lmag);
    return hipGetLastError();
}

hipError_t launch_clock(DevScalars *sc, hipStream_t s) {
    hipLaunchKernelGGL(k_clock, dim3(1), dim3(1), 0, s, sc);
    return hipGetLastError();
}

hipError_t launch_pack(const Params &p, const double *padded, double *field, hipStream_t s) {
    long long nb = (p.fsize + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)nb), dim3(256), 0, s, p, padded, field);
    return hipGetLastError();
}

hipError_t launch_unpack(const Params &p, const double *field, double *padded, hipStream_t s) {
    const long long n = (long long)(p.npx + 2) * (p.npy + 2) * (p.npz + 2);
    long long nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_unpack, dim3((unsigned)nb), dim3(256), 0, s, p, field, padded, 1);
    return hipGetLastError();
}

hipError_t launch_unpack_rhs(const Params &p, const double *field, double *padded, hipStream_t s) {
    const long long n = (long long)(p.npx + 2) * (p.npy + 2) * (p.npz + 2);
    long long nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_unpack, dim3((unsigned)nb), dim3(256), 0, s, p, field, padded, 0);
    return hipGetLastError();
}

}  // namespace rmhd
