// Mixed-precision form of the step (no counterpart in the reference).
//
// The state is carried as FP32 perturbations h about an FP64 hydrostatic
// background (rho0(z), T0(z); zero momentum and field), and the right-hand
// side is evaluated in FP32 in perturbation form: F(bg + h) = F_bg(z) + F'(h),
// where F_bg (nonzero only for the momentum z component and the temperature)
// is the background's own discrete residual, formed once per plane in FP64.
//
// Stored variables: RU, RV, RW, rho', P', BX, BY, BZ, where P' is the
// pressure perturbation rho T - rho0 T0 = rho0 T' + T0 rho' + rho' T'.  In
// low-Mach flow P' is a small residual of rho0 T' and T0 rho'; storing it
// directly keeps the pressure gradient, which drives the horizontal flow,
// accurate to the FP32 precision of P' itself, and T' = (P' - T0 rho')/rho
// follows without cancellation.  (Storing T' instead leaves the horizontal
// momentum several times 1e-6 off the FP64 result after 1000 steps.)
//
// The RK3 accumulation runs in FP64 in registers, on the same variables as
// the FP64 form (T, not P): X = value + D + gamma dt F, where the value of
// the temperature slot is T'(rho', P') as the stencil computes it (trep).  The second
// register D (FP32) carries, per Wray's low-storage scheme, the zeta*dt*F
// increment of the next stage plus the rounding residual of the stored
// value (for T: X_T - T'(new rho', new P')).  After the last stage only the
// residual is left, at most a few FP32 ulps of the value: it is kept in a
// third register R as BF16, two fields per 32-bit word (8 significant bits
// of a residual of relative size 2^-24, i.e. the state to about 2^-33), so
// that (h, R) is a
// compensated pair and the storage rounding does not accumulate from step
// to step.  Stage 1 reads R, stages 1-2 write D, stage 3 writes R.
// The CFL reduction, the boundary planes and the background are FP64.
// The file is compiled without FMA contraction; the FP32 right-hand side
// and the FP64 accumulation of the stage body opt in (#pragma clang fp
// contract), everything that packs, unpacks or reconstructs stored values
// (trep, pprime, the residuals) rounds each operation separately.
//
// Traffic per zone (8 fields): stage 1 reads h (4 B) and R (2 B) and writes
// h and D (4 + 4 B); stage 2 reads and writes h and D; stage 3 reads h and D
// and writes h and R: 112 + 128 + 112 = 352 B/zone/step (FP64: 640).

#include "mhd_device.h"
#include <type_traits>

namespace rmhd {

constexpr int MTX = 64;

// Arithmetic type of the perturbation right-hand side (storage stays FP32).
typedef float rhs_t;

// Tile of TX columns x TY rows, one point per thread (TX = 64: one row per
// wavefront; TX = 32: two rows per wavefront, for narrow ensemble boxes).
template <int TX_, int TY_>
struct MTile {
    static constexpr int TX = TX_;
    static constexpr int TY = TY_;
    static constexpr int NT = TX * TY;
    static constexpr int SW = TX + 2;
    static constexpr int SH = TY + 2;
    static constexpr int SP = SW * SH;
    static constexpr int NHALO = SP - NT;
    static_assert(NHALO <= NT, "one halo point per thread at most");
};

// LDS: raw planes k, k+1 (2 x NF fields) and a three-plane ring of u, v, w,
// T'.
template <int TY, bool MAG, int MODE, int TXW = MTX>
constexpr int mixed_lds_bytes() {
    return (2 * (MAG ? 8 : 5) + 12) * MTile<TXW, TY>::SP * (int)sizeof(float);
}

// P' of temperature perturbation tp at density perturbation rh.  The
// stored P' is formed from the exact (FP64) rho', not from its FP32
// rounding: T0 times the rounding error of rho' would be comparable to P'.
__host__ __device__ __forceinline__ double pprime(double r0, double t0, double rh, double tp) {
    return r0 * tp + t0 * rh + rh * tp;
}

// CFL limits of the mixed form (FP64), accumulated in monotone transformed
// form so that a point costs two divisions and one square root:
//   min dd/sqrt(U)               = 1/sqrt(max U/dd^2)
//   min 0.5 dd^2 REPR CV rho/kap = 0.5 REPR CV min(rho dd^2 / kap)
//   min 0.375 dd^2 RE rho        = 0.375 RE min(rho dd^2)
//   min 0.5 dd^2 RM              = 0.5 RM min(dd^2)
// (U the fast-mode speed squared of sub:750-856).  The workgroup converts
// back once and updates the same keys as cfl_reduce.  NaN sticks.
struct CflFast {
    double mach, q1, s2, s3, d2;
};

__device__ __forceinline__ double nmax(double a, double b) { return (b > a || !(b == b)) ? b : a; }
__device__ __forceinline__ double nmin(double a, double b) { return (b < a || !(b == b)) ? b : a; }

__device__ __forceinline__ CflFast cflf_init() { return CflFast{0.0, 0.0, INFINITY, INFINITY, INFINITY}; }

template <bool MAG>
__device__ __forceinline__ void cflf_point(const Params &p, const double *q, double dd2, double idd2, double irk,
                                           CflFast &c) {
    const double ro = q[F_RO], tt = q[F_TT];
    const double ri = 1.0 / ro;
    const double uu = (q[F_RU] * q[F_RU] + q[F_RV] * q[F_RV] + q[F_RW] * q[F_RW]) * (ri * ri);
    c.mach = nmax(c.mach, p.ogamma * uu / tt);
    double cs = p.gamma * tt;
    if (MAG) cs += (q[F_BX] * q[F_BX] + q[F_BY] * q[F_BY] + q[F_BZ] * q[F_BZ]) * p.obeta * ri;
    const double u2 = uu + cs + 2.0 * sqrt(uu * cs);
    c.q1 = nmax(c.q1, u2 * idd2);
    const double rd = ro * dd2;
    c.s2 = nmin(c.s2, rd * irk);
    c.s3 = nmin(c.s3, rd);
    if (MAG) c.d2 = nmin(c.d2, dd2);
}

// dd^2 = min(ddx, ddy, ddz)^2 and its inverse for one point
struct DdPair {
    double dd2, idd2;
};
__device__ __forceinline__ DdPair ddpair(double ddx, double ddy, double ddz) {
    const double a = ddx * ddx, b = ddy * ddy, c = ddz * ddz;
    const double d2 = fmin(fmin(a, b), c);
    return DdPair{d2, 1.0 / d2};
}

// d2extra: a min(dd^2) contribution computed elsewhere (the grid-only term
// of the stage kernel, which does not reduce it per point).
__device__ void cflf_reduce(const Params &p, CflFast c, DevScalars *__restrict__ sc, double d2extra = INFINITY) {
    __shared__ double red[16][5];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        c.mach = nmax(c.mach, __shfl_xor(c.mach, off, 64));
        c.q1 = nmax(c.q1, __shfl_xor(c.q1, off, 64));
        c.s2 = nmin(c.s2, __shfl_xor(c.s2, off, 64));
        c.s3 = nmin(c.s3, __shfl_xor(c.s3, off, 64));
        c.d2 = nmin(c.d2, __shfl_xor(c.d2, off, 64));
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) {
        red[wv][0] = c.mach; red[wv][1] = c.q1; red[wv][2] = c.s2; red[wv][3] = c.s3; red[wv][4] = c.d2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 1; v < nw; ++v) {
            c.mach = nmax(c.mach, red[v][0]);
            c.q1 = nmax(c.q1, red[v][1]);
            c.s2 = nmin(c.s2, red[v][2]);
            c.s3 = nmin(c.s3, red[v][3]);
            c.d2 = nmin(c.d2, red[v][4]);
        }
        c.d2 = nmin(c.d2, d2extra);
        const double w[4] = {1.0 / sqrt(c.q1), 0.5 * p.repr * p.cv * c.s2, 0.375 * p.re * c.s3, 0.5 * p.rm * c.d2};
        key_raise(&sc->key_mach, dkey(c.mach));
        for (int n = 0; n < 4; ++n) {
            if (n == 3 && !p.lmag) break;
            // a NaN limit becomes -NaN for the min reduction
            key_lower(&sc->key_min[n], (w[n] == w[n]) ? dkey(w[n]) : 0ull);
        }
    }
}

struct RawF {
    float v[NFIELD];
};

template <bool MAG>
__device__ __forceinline__ void load_rawf(const float *__restrict__ q, long long fs, long long off, RawF &r) {
#pragma unroll
    for (int f = 0; f < (MAG ? 8 : 5); ++f) r.v[f] = q[f * fs + off];
}

// One tile column chunk of mixed stage s = MODE (1..3); see stage_tile of
// the FP64 form for the tile organisation.
// Per-plane values (z metrics, background, coefficients) from the
// records of Background::planes: one table pointer for the march.
#define PLM(F, K) uload(&bg.planes[K].F, 0)
#define PLB(F, K) uload(&bg.planes[K].F, 0)
#define PLP(F, K) uload(&bg.planes[K].F, 0)

template <int TY, bool MAG, int MODE, int TXW = MTX>
__device__ __forceinline__ void stage_tile_m(const Params &p, const Metrics &m, const Background &bg,
                                             const float *__restrict__ hin, float *__restrict__ hout,
                                             float *__restrict__ dr, unsigned *__restrict__ rr,
                                             DevScalars *__restrict__ sc, int i0, int j0, int kbeg, int kend,
                                             float *ldsf) {
    using T = MTile<TXW, TY>;
    constexpr int SW = T::SW, SP = T::SP;
    constexpr int NF = MAG ? 8 : 5;
    float *const s_cur = ldsf;
    float *const s_ring = ldsf + 2 * NF * SP;

    const int t = threadIdx.x;
    const int npx = p.npx, npy = p.npy, npz = p.npz;
    const long long fs = p.fsize, pl = p.plane;

    const int olx = t % TXW + 1, oly = t / TXW + 1;
    const int opt = oly * SW + olx;
    const int gi = i0 + olx - 1, gj = j0 + oly - 1;
    const bool valid = (gi < npx) && (gj < npy);
    const long long ooff = (long long)wrapi(gj, npy) * npx + wrapi(gi, npx);
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

    auto cur = [&](int k, int f, int pt) -> float & { return s_cur[((k & 1) * NF + f) * SP + pt]; };
    auto ring = [&](int k, int c, int pt) -> float & { return s_ring[((k % 3) * 4 + c) * SP + pt]; };
    auto put_cur = [&](int k, int pt, const RawF &r) {
#pragma unroll
        for (int f = 0; f < NF; ++f) cur(k, f, pt) = r.v[f];
    };
    auto put_ring = [&](int k, int pt, const RawF &r) {
        const float ri = rcp32(PLB(ro32, k) + r.v[3]);
        ring(k, 0, pt) = r.v[0] * ri;
        ring(k, 1, pt) = r.v[1] * ri;
        ring(k, 2, pt) = r.v[2] * ri;
        ring(k, 3, pt) = trep(PLB(ro32, k), PLB(tt32, k), r.v[3], r.v[4]);  // temperature perturbation
    };

    RawF rm_, rp_o, rp_h;
    load_rawf<MAG>(hin, fs, (long long)(kbeg - 1) * pl + ooff, rm_);
    put_ring(kbeg - 1, opt, rm_);
    if (has_h) {
        load_rawf<MAG>(hin, fs, (long long)(kbeg - 1) * pl + hoff, rp_h);
        put_ring(kbeg - 1, hpt, rp_h);
    }
#pragma unroll 1
    for (int kk = kbeg; kk <= kbeg + 1; ++kk) {
        load_rawf<MAG>(hin, fs, (long long)kk * pl + ooff, rp_o);
        put_cur(kk, opt, rp_o);
        put_ring(kk, opt, rp_o);
        if (has_h) {
            load_rawf<MAG>(hin, fs, (long long)kk * pl + hoff, rp_h);
            put_cur(kk, hpt, rp_h);
            put_ring(kk, hpt, rp_h);
        }
    }

    const rhs_t HX = (rhs_t)p.hx, H2X = (rhs_t)p.h2x, HY = (rhs_t)p.hy, H2Y = (rhs_t)p.h2y;
    const rhs_t HZ = (rhs_t)p.hz, H2Z = (rhs_t)p.h2z;
    const int gix = wrapi(gi, npx) + 1, gjy = wrapi(gj, npy) + 1;
    const rhs_t dx = (rhs_t)m.dxxdx[gix], d2x = (rhs_t)m.d2xxdx2[gix];
    const rhs_t dy = (rhs_t)m.dyydy[gjy], d2y = (rhs_t)m.d2yydy2[gjy];
    const rhs_t hxdx = HX * dx, hydy = HY * dy;
    const rhs_t h2xdx2 = H2X * dx * dx, h2ydy2 = H2Y * dy * dy;
    const rhs_t hxd2x = HX * d2x, hyd2y = HY * d2y;
    const double dt = sc->dt;
    const double coefg = (MODE == 1 ? p.gam1 : MODE == 2 ? p.gam2 : p.gam3) * dt;
    const double coefz = (MODE == 1 ? p.zeta1 : p.zeta2) * dt;
    // stage 3: this column's part of min(dd^2) and its inverse; per plane
    // the CFL takes dd^2 = min(c_d2, ddz^2) and 1/dd^2 = max(c_id2, 1/ddz^2),
    // which is 1/min(...) exactly (correctly rounded division is monotone)
    double c_d2 = 0.0, c_id2 = 0.0;
    if (MODE == 3) {
        const double ddx = m.ddx[gix], ddy = m.ddy[gjy];
        c_d2 = fmin(ddx * ddx, ddy * ddy);
        c_id2 = 1.0 / c_d2;
    }
    // stage 3: this thread's running CFL maxima/minima over its column
    CflFast cfl_acc = cflf_init();
    const bool frc = bg.force != nullptr && *bg.force_on != 0;

    // One plane of the march.  BND instantiates the boundary-plane form
    // (one-sided z stencils, extra planes read directly); the interior form
    // has none of that code, so its loads are waited for only where used.
    auto plane = [&](const int k, auto bnd) __attribute__((always_inline)) {
        constexpr bool BND = decltype(bnd)::value;
        const bool pre = k < kend;
        const long long o = (long long)k * pl + ooff;
        // D (stages 2, 3) or the packed R words (stage 1) of this point, read
        // once; R is unpacked where it is used, so that no wait for these
        // loads comes before the plane's arithmetic.  Every load of the plane
        // is issued on every path (lanes outside the box read their wrapped
        // column, which is in range; the last plane of a chunk and the lanes
        // without a halo point re-read points of their own), D first: the
        // number of loads in flight is then the same on all paths, and the
        // wait for D at the accumulation does not also wait for the prefetch.
        float dk_[8];
        unsigned rk_[4];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            if (MODE != 1) dk_[f] = __builtin_nontemporal_load(&dr[f * fs + o]);
            else if (!(f & 1)) rk_[f >> 1] = __builtin_nontemporal_load(&rr[(f >> 1) * fs + o]);
        }
        {
            const long long kp = (long long)(pre ? k + 2 : k) * pl;
            load_rawf<MAG>(hin, fs, kp + ooff, rp_o);
            load_rawf<MAG>(hin, fs, kp + (has_h ? hoff : ooff), rp_h);
        }
        __syncthreads();
        if (valid) {
            const bool top = BND && k == 1, bot = BND && k == npz;
            const rhs_t C43 = PLP(c43, k), OCV = PLP(ocv, k), ORE = PLP(ore, k), ORM = PLP(orm, k);
            const rhs_t OBETA = PLP(obeta, k), GRAV = PLP(grav, k);
            const rhs_t dz = (rhs_t)PLM(dzzdz, k), d2z = (rhs_t)PLM(d2zzdz2, k);
            const rhs_t rk = (rhs_t)PLM(rkapa, k), dkap = (rhs_t)PLM(dkapa, k);
            const rhs_t hzdz = HZ * dz, hzd2z = HZ * d2z, h2zdz2 = H2Z * dz * dz;
            const rhs_t r0 = PLB(ro32, k), t0 = PLB(tt32, k);
#define CUR(f, d) cur(k, f, opt + (d))
#define NXT(f) cur(k + 1, f, opt)
#define RNG(kp, c, d) ring(kp, c, opt + (d))
            float h0[8];
            rhs_t q0[8];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                h0[f] = CUR(f, 0);
                q0[f] = h0[f];
            }
            const rhs_t rh0 = q0[3], th0 = RNG(k, 3, 0);
            const rhs_t tf0 = t0 + th0;                    // full temperature
            const rhs_t rinv = rcp32(r0 + rh0);                // 1/rho
            const rhs_t u0 = RNG(k, 0, 0), v0 = RNG(k, 1, 0), w0 = RNG(k, 2, 0);
            const rhs_t um = RNG(k - 1, 0, 0), vm = RNG(k - 1, 1, 0), wm = RNG(k - 1, 2, 0);
            const rhs_t up = RNG(k + 1, 0, 0), vp = RNG(k + 1, 1, 0), wp = RNG(k + 1, 2, 0);
            const rhs_t uxp = RNG(k, 0, 1), uxm = RNG(k, 0, -1), uyp = RNG(k, 0, SW), uym = RNG(k, 0, -SW);
            const rhs_t vxp = RNG(k, 1, 1), vxm = RNG(k, 1, -1), vyp = RNG(k, 1, SW), vym = RNG(k, 1, -SW);
            const rhs_t wxp = RNG(k, 2, 1), wxm = RNG(k, 2, -1), wyp = RNG(k, 2, SW), wym = RNG(k, 2, -SW);
            const rhs_t ru_m = rm_.v[0], rv_m = rm_.v[1], rw_m = rm_.v[2];
            const rhs_t ru_p = NXT(0), rv_p = NXT(1), rw_p = NXT(2);
            const rhs_t th_m = RNG(k - 1, 3, 0), th_p = RNG(k + 1, 3, 0);
            // stored pressure perturbation at the six neighbours
            const rhs_t pm = rm_.v[4], pp = NXT(4);
            const rhs_t pxm = CUR(4, -1), pxp = CUR(4, 1), pym = CUR(4, -SW), pyp = CUR(4, SW);

            rhs_t rwk2 = rhs_t(0), wk2 = rhs_t(0), bxk2 = rhs_t(0), bxk3 = rhs_t(0), byk2 = rhs_t(0), byk3 = rhs_t(0);
            if (top || bot) {
                const int ka = top ? k + 2 : k - 2;
                const int kb = top ? k + 3 : k - 3;
                const long long oa = (long long)ka * pl + ooff, ob = (long long)kb * pl + ooff;
                rwk2 = hin[2 * fs + oa];
                wk2 = rwk2 * rcp32((rhs_t)PLB(ro32, ka) + (rhs_t)hin[3 * fs + oa]);
                if (MAG) {
                    bxk2 = hin[5 * fs + oa];
                    byk2 = hin[6 * fs + oa];
                    bxk3 = hin[5 * fs + ob];
                    byk3 = hin[6 * fs + ob];
                }
            }

            // The FP32 right-hand side is evaluated with its multiply-adds
            // fused (one rounding instead of two).  The mixed form is judged
            // by its error against the FP64 form, not bit for bit; fusion is
            // decided per source expression, so every kernel that inlines
            // this body computes the same values.
            rhs_t fr, fw, fv, fu, ft, w1 = rhs_t(0), w2 = rhs_t(0), w3 = rhs_t(0);
            {
#pragma clang fp contract(on)
                // continuity
                fr = (CUR(0, -1) - CUR(0, 1)) * hxdx - (CUR(1, SW) - CUR(1, -SW)) * hydy;
                if (top) fr -= (rhs_t(4) * rw_p - rwk2) * hzdz;
                else if (bot) fr -= (rwk2 - rhs_t(4) * rw_m) * hzdz;
                else fr -= (rw_p - rw_m) * hzdz;
                // buoyancy and pressure gradient of the perturbation
                fw = GRAV * rh0 - (pp - pm) * hzdz;
                fv = (pym - pyp) * hydy;
                fu = (pxm - pxp) * hxdx;
                // momentum advection
                fu -= (CUR(0, 1) * uxp - CUR(0, -1) * uxm) * hxdx + (CUR(0, SW) * vyp - CUR(0, -SW) * vym) * hydy
                      + (ru_p * wp - ru_m * wm) * hzdz;
                fv -= (CUR(1, 1) * uxp - CUR(1, -1) * uxm) * hxdx + (CUR(1, SW) * vyp - CUR(1, -SW) * vym) * hydy
                      + (rv_p * wp - rv_m * wm) * hzdz;
                fw -= (CUR(2, 1) * uxp - CUR(2, -1) * uxm) * hxdx + (CUR(2, SW) * vyp - CUR(2, -SW) * vym) * hydy
                      + (rw_p * wp - rw_m * wm) * hzdz;
                // temperature advection, the background gradient included
                const rhs_t txp = RNG(k, 3, 1), txm = RNG(k, 3, -1), typ = RNG(k, 3, SW), tym = RNG(k, 3, -SW);
                const rhs_t dtz = (th_p - th_m) * hzdz;
                ft = -(u0 * (txp - txm) * hxdx + v0 * (typ - tym) * hydy + w0 * (dtz + PLB(dt0dz32, k)));
                // conduction of T': (1/rho) c [kappa lap T' + kappa' dT'/dz]; plus the
                // change of the background's conduction with 1/rho: -rho'/(rho rho0) S0
                {
                    const rhs_t lap = (txp - txm) * hxd2x + (txp - rhs_t(2) * th0 + txm) * h2xdx2
                                      + (typ - tym) * hyd2y + (typ - rhs_t(2) * th0 + tym) * h2ydy2
                                      + (th_p - th_m) * hzd2z + (th_p - rhs_t(2) * th0 + th_m) * h2zdz2;
                    ft += rinv * PLP(t_cond, k) * (rk * lap + dkap * dtz) - rh0 * rinv * PLB(rinv0_32, k) * PLB(s0_32, k);
                }
                // viscous Laplacians
                fu += ORE * (C43 * ((uxp - uxm) * hxd2x + (uxp - rhs_t(2) * u0 + uxm) * h2xdx2)
                             + (uyp - uym) * hyd2y + (uyp - rhs_t(2) * u0 + uym) * h2ydy2
                             + (up - um) * hzd2z + (up - rhs_t(2) * u0 + um) * h2zdz2);
                fv += ORE * ((vxp - vxm) * hxd2x + (vxp - rhs_t(2) * v0 + vxm) * h2xdx2
                             + C43 * ((vyp - vym) * hyd2y + (vyp - rhs_t(2) * v0 + vym) * h2ydy2)
                             + (vp - vm) * hzd2z + (vp - rhs_t(2) * v0 + vm) * h2zdz2);
                fw += ORE * ((wxp - wxm) * hxd2x + (wxp - rhs_t(2) * w0 + wxm) * h2xdx2
                             + (wyp - wym) * hyd2y + (wyp - rhs_t(2) * w0 + wym) * h2ydy2
                             + C43 * ((wp - wm) * hzd2z + (wp - rhs_t(2) * w0 + wm) * h2zdz2));
                // heating and grad div u cross terms
                const rhs_t tv = PLP(t_visc, k), tv43 = PLP(t_visc43, k), tcr = PLP(t_cross, k);
                const rhs_t dxu = (uxp - uxm) * hxdx, dxv = (vxp - vxm) * hxdx, dxw = (wxp - wxm) * hxdx;
                const rhs_t dyu = (uyp - uym) * hydy, dyv = (vyp - vym) * hydy, dyw = (wyp - wym) * hydy;
                const rhs_t dzu = (up - um) * hzdz, dzv = (vp - vm) * hzdz, dzw = (wp - wm) * hzdz;
                ft += (C43 * dxu * dxu + dxv * dxv + dxw * dxw + dyu * dyu + rhs_t(2) * dyu * dxv) * rinv * tv
                      + (dyv * dyv + dzw * dzw - dxu * dzw - dxu * dyv - dyv * dzw) * rinv * tv43
                      + (dzu * dzu + dzv * dzv + dyw * dyw + rhs_t(2) * dzv * dyw + rhs_t(2) * dzu * dxw) * rinv * tv
                      - OCV * (dxu + dyv + dzw) * tf0;
                fu += tcr * (((RNG(k, 1, SW + 1) - RNG(k, 1, SW - 1)) - (RNG(k, 1, -SW + 1) - RNG(k, 1, -SW - 1))) * hxdx * hydy
                             + ((RNG(k + 1, 2, 1) - RNG(k - 1, 2, 1)) - (RNG(k + 1, 2, -1) - RNG(k - 1, 2, -1))) * hzdz * hxdx);
                fv += tcr * (((RNG(k, 0, SW + 1) - RNG(k, 0, SW - 1)) - (RNG(k, 0, -SW + 1) - RNG(k, 0, -SW - 1))) * hxdx * hydy
                             + ((RNG(k + 1, 2, SW) - RNG(k - 1, 2, SW)) - (RNG(k + 1, 2, -SW) - RNG(k - 1, 2, -SW))) * hzdz * hydy);
                fw += tcr * (((RNG(k + 1, 0, 1) - RNG(k - 1, 0, 1)) - (RNG(k + 1, 0, -1) - RNG(k - 1, 0, -1))) * hzdz * hxdx
                             + ((RNG(k + 1, 1, SW) - RNG(k - 1, 1, SW)) - (RNG(k + 1, 1, -SW) - RNG(k - 1, 1, -SW))) * hzdz * hydy);
                // Coriolis
                if (p.lrot) {
                    const rhs_t omx = PLP(omx, k), omz = PLP(omz, k);
                    fu += omz * q0[1];
                    fv += -omz * q0[0] + omx * q0[2];
                    fw -= omx * q0[1];
                }
                // MHD
                if (MAG) {
                    const rhs_t bx0 = q0[5], by0 = q0[6], bz0 = q0[7];
                    const rhs_t bxm = rm_.v[5], bym = rm_.v[6], bzm = rm_.v[7];
                    const rhs_t bxp = NXT(5), byp = NXT(6), bzp = NXT(7);
                    const rhs_t bxxp = CUR(5, 1), bxxm = CUR(5, -1), bxyp = CUR(5, SW), bxym = CUR(5, -SW);
                    const rhs_t byxp = CUR(6, 1), byxm = CUR(6, -1), byyp = CUR(6, SW), byym = CUR(6, -SW);
                    const rhs_t bzxp = CUR(7, 1), bzxm = CUR(7, -1), bzyp = CUR(7, SW), bzym = CUR(7, -SW);
                    rhs_t dzw1 = dzw;
                    if (top) dzw1 = (rhs_t(4) * wp - wk2) * hzdz;
                    else if (bot) dzw1 = (wk2 - rhs_t(4) * wm) * hzdz;
                    const rhs_t dxbx = (bxxp - bxxm) * hxdx, dxby = (byxp - byxm) * hxdx, dxbz = (bzxp - bzxm) * hxdx;
                    const rhs_t dybx = (bxyp - bxym) * hydy, dyby = (byyp - byym) * hydy, dybz = (bzyp - bzym) * hydy;
                    const rhs_t dzbx = (bxp - bxm) * hzdz, dzby = (byp - bym) * hzdz, dzbz = (bzp - bzm) * hzdz;
                    // induction (B.grad) u - B div u - (u.grad) B
                    w1 = by0 * dyu - bx0 * dyv + bz0 * dzu - bx0 * dzw1 - u0 * dxbx - v0 * dybx - w0 * dzbx;
                    w2 = bx0 * dxv - by0 * dxu + bz0 * dzv - by0 * dzw1 - u0 * dxby - v0 * dyby - w0 * dzby;
                    w3 = bx0 * dxw - bz0 * dxu + by0 * dyw - bz0 * dyv - u0 * dxbz - v0 * dybz - w0 * dzbz;
                    // Lorentz force
                    fu += OBETA * (by0 * (dybx - dxby) - bz0 * dxbz + bz0 * dzbx);
                    fv += OBETA * (bx0 * (dxby - dybx) + bz0 * (dzby - dybz));
                    fw += OBETA * (bx0 * dxbz - bx0 * dzbx - by0 * (dzby - dybz));
                    // Ohmic heating |J|^2 / rho
                    ft += ((dybx - dxby) * (dybx - dxby) + dxbz * dxbz + dzbx * dzbx - rhs_t(2) * dzbx * dxbz
                           + (dzby - dybz) * (dzby - dybz)) * rinv * PLP(t_ohm, k);
                    // resistive diffusion, one-sided in z on the boundary planes for BX, BY
                    rhs_t d2zbx, d2zby;
                    if (top) {
                        d2zbx = (-rhs_t(3) * bx0 + rhs_t(4) * bxp - bxk2) * hzd2z + (rhs_t(2) * bx0 - rhs_t(5) * bxp + rhs_t(4) * bxk2 - bxk3) * h2zdz2;
                        d2zby = (-rhs_t(3) * by0 + rhs_t(4) * byp - byk2) * hzd2z + (rhs_t(2) * by0 - rhs_t(5) * byp + rhs_t(4) * byk2 - byk3) * h2zdz2;
                    } else if (bot) {
                        d2zbx = (rhs_t(3) * bx0 - rhs_t(4) * bxm + bxk2) * hzd2z + (rhs_t(2) * bx0 - rhs_t(5) * bxm + rhs_t(4) * bxk2 - bxk3) * h2zdz2;
                        d2zby = (rhs_t(3) * by0 - rhs_t(4) * bym + byk2) * hzd2z + (rhs_t(2) * by0 - rhs_t(5) * bym + rhs_t(4) * byk2 - byk3) * h2zdz2;
                    } else {
                        d2zbx = (bxp - bxm) * hzd2z + (bxp - rhs_t(2) * bx0 + bxm) * h2zdz2;
                        d2zby = (byp - bym) * hzd2z + (byp - rhs_t(2) * by0 + bym) * h2zdz2;
                    }
                    const rhs_t d2zbz = (bzp - bzm) * hzd2z + (bzp - rhs_t(2) * bz0 + bzm) * h2zdz2;
                    w1 += ORM * ((bxxp - bxxm) * hxd2x + (bxxp - rhs_t(2) * bx0 + bxxm) * h2xdx2
                                 + (bxyp - bxym) * hyd2y + (bxyp - rhs_t(2) * bx0 + bxym) * h2ydy2 + d2zbx);
                    w2 += ORM * ((byxp - byxm) * hxd2x + (byxp - rhs_t(2) * by0 + byxm) * h2xdx2
                                 + (byyp - byym) * hyd2y + (byyp - rhs_t(2) * by0 + byym) * h2ydy2 + d2zby);
                    w3 += ORM * ((bzxp - bzxm) * hxd2x + (bzxp - rhs_t(2) * bz0 + bzxm) * h2xdx2
                                 + (bzyp - bzym) * hyd2y + (bzyp - rhs_t(2) * bz0 + bzym) * h2ydy2 + d2zbz);
                }
#undef CUR
#undef NXT
#undef RNG
            }
            // FP64 accumulation: F = F_bg(z) + F', X = value + D + gamma dt F
            // (the temperature slot's value is T'(rho', P')); store the new
            // FP32 values and carry the residual (+ zeta dt F) in D.  The
            // products dt F are fused into their sums; the residual itself
            // (X minus its stored representation) is exact either way.
            double Fd[8] = {fu, fv, fw, fr, ft, w1, w2, w3};
            Fd[2] += PLB(fwbg, k);
            Fd[4] += PLB(ftbg, k);
            if (frc) {
                const float *g = bg.force + o;
                Fd[0] += (double)g[0];
                Fd[1] += (double)g[fs];
                Fd[2] += (double)g[2 * fs];
                Fd[4] += (double)g[3 * fs];
                if (MAG) {
                    Fd[5] += (double)g[4 * fs];
                    Fd[6] += (double)g[5 * fs];
                    Fd[7] += (double)g[6 * fs];
                }
            }
            const double r0d = PLB(ro, k), t0d = PLB(tt, k);
            double xq[8];
            float hn[8];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
#pragma clang fp contract(on)
                const double v = f == 4 ? (double)th0 : (double)h0[f];
                // (read only once this field's right-hand side exists)
                const float dkf = MODE == 1 ? r_half(after(rk_[f >> 1], Fd[f]), f) : after(dk_[f], Fd[f]);
                xq[f] = (v + (double)dkf) + coefg * Fd[f];
                if (f != 4) hn[f] = (float)xq[f];
            }
            hn[4] = (float)pprime(r0d, t0d, xq[3], xq[4]);
            float rs[8];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const double r = xq[f] - (f == 4 ? (double)trep(PLB(ro32, k), PLB(tt32, k), hn[3], hn[4]) : (double)hn[f]);
                __builtin_nontemporal_store(hn[f], &hout[f * fs + o]);
                if (MODE < 3) {
#pragma clang fp contract(on)
                    __builtin_nontemporal_store((float)(r + coefz * Fd[f]), &dr[f * fs + o]);
                } else {
                    rs[f] = (float)r;
                }
            }
            if (MODE == 3) {
                // the residuals as BF16 pairs (an odd last field with a zero high half)
#pragma unroll
                for (int w = 0; w < (NF + 1) / 2; ++w)
                    __builtin_nontemporal_store(pack_bf16x2(rs[2 * w], 2 * w + 1 < NF ? rs[2 * w + 1] : 0.0f),
                                                &rr[w * fs + o]);
            }
            if (MODE == 3 && !top && !bot) {
                double full[8];
#pragma unroll
                for (int f = 0; f < NF; ++f) full[f] = xq[f];
                full[3] += PLB(ro, k);
                full[4] += PLB(tt, k);
                const double ddz = PLM(ddz, k);
                const double dd2 = fmin(c_d2, ddz * ddz), idd2 = fmax(c_id2, PLB(iddz2, k));
                // (min(dd^2) is not accumulated per point: see the reduction)
                CflFast c{cfl_acc.mach, cfl_acc.q1, cfl_acc.s2, cfl_acc.s3, INFINITY};
                cflf_point<MAG>(p, full, dd2, idd2, PLB(irk, k), c);
                cfl_acc = c;
            }
#pragma unroll
            for (int f = 0; f < NF; ++f) rm_.v[f] = h0[f];
        }
        __syncthreads();
        if (pre) {
            put_cur(k + 2, opt, rp_o);
            put_ring(k + 2, opt, rp_o);
            if (has_h) {
                put_cur(k + 2, hpt, rp_h);
                put_ring(k + 2, hpt, rp_h);
            }
        }
    };
#pragma unroll 1
    for (int k = kbeg; k <= kend; ++k) {
        if (k == 1 || k == npz) plane(k, std::true_type{});
        else plane(k, std::false_type{});
    }
    if (MODE == 3) {
        // min(dd^2) depends on the grid only: its minimum over all points
        // comes with the metrics
        cflf_reduce(p, CflFast{cfl_acc.mach, cfl_acc.q1, cfl_acc.s2, cfl_acc.s3, INFINITY}, sc,
                    MAG ? sc->dd2min : INFINITY);
    }
}

template <int TY, bool MAG, int MODE>
__global__ __launch_bounds__(MTX *TY, 3) void k_stage_m(const Params p, const Metrics m, const Background bg,
                                                     const float *__restrict__ hin, float *__restrict__ hout,
                                                     float *__restrict__ dr, unsigned *__restrict__ rr,
                                                     DevScalars *__restrict__ sc, int kchunk) {
    extern __shared__ float ldsf[];
    const int ntx = (p.npx + MTX - 1) / MTX, nty = (p.npy + TY - 1) / TY;
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int tz = bid / (ntx * nty);
    const int rem = bid - tz * (ntx * nty);
    const int tyb = rem / ntx, txb = rem - tyb * ntx;
    const int kbeg = 1 + tz * kchunk;
    const int kend = min(p.npz, kbeg + kchunk - 1);
    if (kbeg > p.npz) return;
    stage_tile_m<TY, MAG, MODE>(p, m, bg, hin, hout, dr, rr, sc, txb * MTX, tyb * TY, kbeg, kend, ldsf);
}

// Ensemble: persistent grid over the work items of all boxes (static stride);
// items are TXW x TY tiles.
template <int TY, bool MAG, int MODE, int TXW>
__global__ __launch_bounds__(TXW *TY, 3) void k_stage_m_batch(const BoxDev *__restrict__ boxes,
                                                           const StageItem *__restrict__ items, int nitems, int a) {
    extern __shared__ float ldsf[];
    const int slot = xcd_remap(blockIdx.x, gridDim.x);
    for (int w = slot; w < nitems; w += gridDim.x) {
        const StageItem it = items[w];
        const BoxDev &B = boxes[it.box];
        __syncthreads();  // the previous item's last LDS reads are done
        stage_tile_m<TY, MAG, MODE, TXW>(B.p, B.m, B.bg, B.h[a], B.h[1 - a], B.dr, B.rr, B.sc, it.i0, it.j0, it.kbeg,
                                    it.kend, ldsf);
    }
}

// BCON of the mixed form, in FP64 on the full state bg + value (+ R after
// the last stage, where R holds the rounding residual).  New boundary
// values are split again into stored values and the residual (D after
// stages 1-2, R after stage 3); a new boundary temperature is stored as the
// P' of the plane's stored rho'.
template <bool MAG, bool LAST>
__device__ __forceinline__ void bcon_column_m(const Params &p, const Metrics &m, const Background &bg,
                                              float *__restrict__ h, float *__restrict__ dr,
                                              unsigned *__restrict__ rr, DevScalars *__restrict__ sc,
                                              int idx) {
    CflFast cfl = cflf_init();
    if (idx < p.npx * p.npy) {
        const long long fs = p.fsize, pl = p.plane;
        const int nz = p.npz;
        const double C13 = p.c13, C23 = p.c23, C43 = p.c43;
        auto full = [&](int f, int k) -> double {
            const long long o = f * fs + (long long)k * pl + idx;
            double v = f == 4 ? (double)trep(bg.ro32[k], bg.tt32[k], h[o - fs], h[o]) : (double)h[o];
            if (LAST) v += (double)r_get(rr, fs, f, (long long)k * pl + idx);
            if (f == 3) v += bg.ro[k];
            if (f == 4) v += bg.tt[k];
            return v;
        };
        auto residual = [&](int f, int k, double r) {
            if (LAST) r_put(rr, fs, f, (long long)k * pl + idx, (float)r);
            else dr[f * fs + (long long)k * pl + idx] = (float)r;
        };
        auto put = [&](int f, int k, double v) {
            const long long o = f * fs + (long long)k * pl + idx;
            if (f == 4) {
                const double tp = v - bg.tt[k];
                const float rh = h[o - fs];
                const double rhx = LAST ? (double)rh + (double)r_get(rr, fs, 3, (long long)k * pl + idx) : (double)rh;
                const float hp = (float)pprime(bg.ro[k], bg.tt[k], rhx, tp);
                h[o] = hp;
                residual(4, k, tp - (double)trep(bg.ro32[k], bg.tt32[k], rh, hp));
                return;
            }
            if (f == 3) v -= bg.ro[k];
            const float hn = (float)v;
            h[o] = hn;
            residual(f, k, v - (double)hn);
        };
        for (int side = 0; side < 2; ++side) {
            const int a = side ? nz : 1, b = side ? nz - 1 : 2, c = side ? nz - 2 : 3;
            const double roa = full(3, a), rob = full(3, b), roc = full(3, c);
            put(0, a, roa * (C43 * full(0, b) / rob - C13 * full(0, c) / roc));
            put(1, a, roa * (C43 * full(1, b) / rob - C13 * full(1, c) / roc));
            put(2, a, 0.0);
            const int code = side ? p.izc : p.itc;
            double tnew;
            if (code == 0) {
                tnew = side ? p.tb : 1.0;
            } else {
                const double dz = p.dzf / m.dzzdz[a];
                tnew = side ? C43 * full(4, b) - C13 * full(4, c) + C23 * dz * p.theta / m.rkapa[a]
                            : C43 * full(4, b) - C13 * full(4, c) - C23 * dz * p.theta;
            }
            put(4, a, tnew);
            if (MAG && p.ibc == 0) put(7, a, 0.0);
        }
        if (LAST) {
            const int j = idx / p.npx, i = idx - j * p.npx;
            for (int side = 0; side < 2; ++side) {
                const int k = side ? nz : 1;
                double v[8];
                for (int f = 0; f < (MAG ? 8 : 5); ++f) v[f] = full(f, k);
                const DdPair d = ddpair(m.ddx[i + 1], m.ddy[j + 1], m.ddz[k]);
                cflf_point<MAG>(p, v, d.dd2, d.idd2, 1.0 / m.rkapa[k], cfl);
            }
        }
    }
    if (LAST) cflf_reduce(p, cfl, sc);
}

template <bool MAG, bool LAST>
__global__ __launch_bounds__(256) void k_bcon_m(const Params p, const Metrics m, const Background bg,
                                                float *__restrict__ h, float *__restrict__ dr,
                                                unsigned *__restrict__ rr, DevScalars *__restrict__ sc) {
    bcon_column_m<MAG, LAST>(p, m, bg, h, dr, rr, sc, blockIdx.x * blockDim.x + threadIdx.x);
}

// Ensemble: one workgroup per (box, run of 256 columns) of register 1-a.
template <bool MAG, bool LAST>
__global__ __launch_bounds__(256) void k_bcon_m_batch(const BoxDev *__restrict__ boxes,
                                                      const ColumnItem *__restrict__ items, int a) {
    const ColumnItem it = items[blockIdx.x];
    const BoxDev &B = boxes[it.box];
    bcon_column_m<MAG, LAST>(B.p, B.m, B.bg, B.h[1 - a], B.dr, B.rr, B.sc, it.col0 + threadIdx.x);
}

// Stand-alone CFL of the mixed state (first step after an upload).
template <bool MAG>
__global__ __launch_bounds__(256) void k_cfl_m(const Params p, const Metrics m, const Background bg,
                                               const float *__restrict__ h, const unsigned *__restrict__ rr,
                                               DevScalars *__restrict__ sc) {
    const long long fs = p.fsize, pl = p.plane;
    const long long n = (long long)p.npz * pl;
    CflFast cfl = cflf_init();
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long o = pl + idx;
        const int k = (int)(idx / pl) + 1;
        const long long r2 = idx - (long long)(k - 1) * pl;
        const int j = (int)(r2 / p.npx), i = (int)(r2 - (long long)j * p.npx);
        double v[8];
        for (int f = 0; f < (MAG ? 8 : 5); ++f) {
            const double x = f == 4 ? (double)trep(bg.ro32[k], bg.tt32[k], h[3 * fs + o], h[4 * fs + o]) : (double)h[f * fs + o];
            v[f] = x + (double)r_get(rr, fs, f, o);
        }
        v[3] += bg.ro[k];
        v[4] += bg.tt[k];
        const DdPair d = ddpair(m.ddx[i + 1], m.ddy[j + 1], m.ddz[k]);
        cflf_point<MAG>(p, v, d.dd2, d.idd2, 1.0 / m.rkapa[k], cfl);
    }
    cflf_reduce(p, cfl, sc);
}

// padded FP64 field f -> (stored value, R = rounding residual) in the
// registers hreg, rreg (all fields); the temperature is stored as the P' of
// the already packed rho' of the same register.
__global__ void k_pack_m(const Params p, const double *__restrict__ padded, int field, const Background bg,
                         float *__restrict__ hreg, unsigned *__restrict__ rreg) {
    const long long n = p.plane * (p.npz + 2), fs = p.fsize;  // n: the field's values
    const int nxp = p.npx + 2, nyp = p.npy + 2;
    float *const h = hreg + field * fs;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long k = idx / p.plane;
        const long long r = idx - k * p.plane;
        const int j = (int)(r / p.npx), i = (int)(r - (long long)j * p.npx);
        double v = padded[(k * nyp + (j + 1)) * nxp + (i + 1)];
        if (field == 4) {
            const double tp = v - bg.tt[k];
            const float rh = hreg[3 * fs + idx];
            const float hp = (float)pprime(bg.ro[k], bg.tt[k], (double)rh + (double)r_get(rreg, fs, 3, idx), tp);
            h[idx] = hp;
            r_put(rreg, fs, 4, idx, (float)(tp - (double)trep(bg.ro32[k], bg.tt32[k], rh, hp)));
            continue;
        }
        if (field == 3) v -= bg.ro[k];
        const float hn = (float)v;
        h[idx] = hn;
        r_put(rreg, fs, field, idx, (float)(v - (double)hn));
    }
}

// field f of (stored value, R) -> padded FP64 field with periodic ghosts
__global__ void k_unpack_m(const Params p, const float *__restrict__ hreg, const unsigned *__restrict__ rreg,
                           int field, const Background bg, double *__restrict__ padded) {
    const int nxp = p.npx + 2, nyp = p.npy + 2, nzp = p.npz + 2;
    const long long n = (long long)nxp * nyp * nzp, fs = p.fsize;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        const long long k = idx / ((long long)nxp * nyp);
        const long long r = idx - k * nxp * nyp;
        const int J = (int)(r / nxp), I = (int)(r - (long long)J * nxp);
        const long long o = k * p.plane + (long long)wrapi(J - 1, p.npy) * p.npx + wrapi(I - 1, p.npx);
        double v;
        if (field == 4) v = bg.tt[k] + (double)trep(bg.ro32[k], bg.tt32[k], hreg[3 * fs + o], hreg[4 * fs + o]);
        else v = (double)hreg[field * fs + o] + (field == 3 ? bg.ro[k] : 0.0);
        padded[idx] = v + (double)r_get(rreg, fs, field, o);
    }
}

// Re-centring of the background (mhd_rebase): rho and T of every interior
// point, reconstructed about the old background ob exactly as k_unpack_m
// does, are split about the new background nb exactly as k_pack_m does
// (rho' before T', whose stored P' uses the new rho' + R).  The other fields
// do not depend on the background; the z ghost planes keep their background.
__global__ void k_rebase_m(const Params p, const Background ob, const Background nb, float *__restrict__ hreg,
                           unsigned *__restrict__ rreg) {
    const long long fs = p.fsize, n = (long long)p.npz * p.plane;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x) {
        const long long idx = p.plane + v;
        const int k = (int)(idx / p.plane);
        const float h3 = hreg[3 * fs + idx], h4 = hreg[4 * fs + idx];
        const double ro = ((double)h3 + ob.ro[k]) + (double)r_get(rreg, fs, 3, idx);
        const double tt = (ob.tt[k] + (double)trep(ob.ro32[k], ob.tt32[k], h3, h4)) + (double)r_get(rreg, fs, 4, idx);
        const double rp = ro - nb.ro[k];
        const float n3 = (float)rp;
        hreg[3 * fs + idx] = n3;
        r_put(rreg, fs, 3, idx, (float)(rp - (double)n3));
        const double tp = tt - nb.tt[k];
        const float n4 = (float)pprime(nb.ro[k], nb.tt[k], (double)n3 + (double)r_get(rreg, fs, 3, idx), tp);
        hreg[4 * fs + idx] = n4;
        r_put(rreg, fs, 4, idx, (float)(tp - (double)trep(nb.ro32[k], nb.tt32[k], n3, n4)));
    }
}

// Network inputs of the closure from the state (h, R) about the background.
__global__ void k_closure_inputs(const Params p, const Background bg, const float *__restrict__ h,
                                 const unsigned *__restrict__ rr, const float *__restrict__ nrm,
                                 __bf16 *__restrict__ x) {
    const long long fs = p.fsize, n = (long long)p.npz * p.plane;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(v / p.plane) + 1;
        const long long o = p.plane + v;
        auto val = [&](int f) { return (double)h[f * fs + o] + (double)r_get(rr, fs, f, o); };
        const double ro = bg.ro[k] + val(3);
        const double tt = bg.tt[k] + (double)trep(bg.ro32[k], bg.tt32[k], h[3 * fs + o], h[4 * fs + o]) +
                          (double)r_get(rr, fs, 4, o);
        const double ri = 1.0 / ro;
        double c[8];
        c[0] = ro;
        c[1] = val(0) * ri;
        c[2] = val(1) * ri;
        c[3] = val(2) * ri;
        c[4] = log(tt) / (p.gamma - 1.0) - log(ro);
        for (int f = 5; f < 8; ++f) c[f] = p.lmag ? val(f) : 0.0;
        // nrm: per-plane means [8][npz], then 1/std [8]
        for (int f = 0; f < 8; ++f)
            x[v * 8 + f] = (__bf16)(float)((c[f] - (double)nrm[f * p.npz + k - 1]) * (double)nrm[8 * p.npz + f]);
    }
}

__global__ void k_set_flag(int *flag) { *flag = 1; }

// Forcing fields from the network output y[v][8] (7 channels used).
__global__ void k_closure_forcing(const Params p, const float *__restrict__ y, const float *__restrict__ nrm,
                                  float *__restrict__ force, int *__restrict__ force_on) {
    const long long fs = p.fsize, n = (long long)p.npz * p.plane;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n;
         v += (long long)gridDim.x * blockDim.x) {
        const long long o = p.plane + v;
        const int k = (int)(v / p.plane);  // plane index 0..npz-1
        bool ok = true;
        // nrm: per-plane mean subgrid terms [7][npz], then std [7]
        for (int c = 0; c < 7; ++c) {
            const float g = y[v * 8 + c] * nrm[7 * p.npz + c] + nrm[c * p.npz + k];
            ok = ok && (g - g == 0.0f);
            force[c * fs + o] = g;
        }
        if (!ok) *force_on = 0;
    }
}

// ---------------------------------------------------------------------------

// z chunks of about 32 planes (the tiles of a co-resident cohort stay closer
// in step over a short march: 512x512x256, 8 chunks 1% faster than 4), at
// least 4096 workgroups, at least 8 planes.
static int chunk_m(const Params &p, int ty, int *nchunks) {
    const int ntiles = ((p.npx + MTX - 1) / MTX) * ((p.npy + ty - 1) / ty);
    int nc = (p.npz + 31) / 32;
    const int need = (4096 + ntiles - 1) / ntiles;
    if (nc < need) nc = need;
    int maxc = p.npz / 8;
    if (maxc < 1) maxc = 1;
    if (nc > maxc) nc = maxc;
    if (nc < 1) nc = 1;
    const int kc = (p.npz + nc - 1) / nc;
    *nchunks = (p.npz + kc - 1) / kc;
    return kc;
}

template <int TY, bool MAG, int MODE>
static hipError_t launch_m(const Params &p, const Metrics &m, const Background &bg, const float *hin, float *hout,
                           float *dr, unsigned *rr, DevScalars *sc, hipStream_t s) {
    constexpr int lds = mixed_lds_bytes<TY, MAG, MODE>();
    static std::atomic<unsigned long long> attr{0};
    {
        hipError_t e = lds_limit_once(attr, reinterpret_cast<const void *>(&k_stage_m<TY, MAG, MODE>), lds);
        if (e != hipSuccess) return e;
    }
    int nch;
    const int kc = chunk_m(p, TY, &nch);
    const int ntiles = ((p.npx + MTX - 1) / MTX) * ((p.npy + TY - 1) / TY);
    hipLaunchKernelGGL((k_stage_m<TY, MAG, MODE>), dim3(ntiles * nch), dim3(MTX * TY), lds, s, p, m, bg, hin, hout,
                       dr, rr, sc, kc);
    return hipGetLastError();
}

template <int TY>
static hipError_t launch_m_ty(int mode, const Params &p, const Metrics &m, const Background &bg, const float *hin,
                              float *hout, float *dr, unsigned *rr, DevScalars *sc, hipStream_t s) {
    if (p.lmag) {
        if (mode == 1) return launch_m<TY, true, 1>(p, m, bg, hin, hout, dr, rr, sc, s);
        if (mode == 2) return launch_m<TY, true, 2>(p, m, bg, hin, hout, dr, rr, sc, s);
        return launch_m<TY, true, 3>(p, m, bg, hin, hout, dr, rr, sc, s);
    }
    if (mode == 1) return launch_m<TY, false, 1>(p, m, bg, hin, hout, dr, rr, sc, s);
    if (mode == 2) return launch_m<TY, false, 2>(p, m, bg, hin, hout, dr, rr, sc, s);
    return launch_m<TY, false, 3>(p, m, bg, hin, hout, dr, rr, sc, s);
}

hipError_t launch_stage_mixed(int mode, const Params &p, const Metrics &m, const Background &bg, const float *hin,
                              float *hout, float *dr, unsigned *rr, DevScalars *sc, hipStream_t s) {
    if (p.tile_y == 8) return launch_m_ty<8>(mode, p, m, bg, hin, hout, dr, rr, sc, s);
    return launch_m_ty<4>(mode, p, m, bg, hin, hout, dr, rr, sc, s);
}

hipError_t launch_bcon_mixed(const Params &p, const Metrics &m, const Background &bg, float *h, float *dr,
                             unsigned *rr, DevScalars *sc, bool last, hipStream_t s) {
    const int n = p.npx * p.npy;
    dim3 grid((n + 255) / 256), block(256);
    if (p.lmag) {
        if (last) hipLaunchKernelGGL((k_bcon_m<true, true>), grid, block, 0, s, p, m, bg, h, dr, rr, sc);
        else hipLaunchKernelGGL((k_bcon_m<true, false>), grid, block, 0, s, p, m, bg, h, dr, rr, sc);
    } else {
        if (last) hipLaunchKernelGGL((k_bcon_m<false, true>), grid, block, 0, s, p, m, bg, h, dr, rr, sc);
        else hipLaunchKernelGGL((k_bcon_m<false, false>), grid, block, 0, s, p, m, bg, h, dr, rr, sc);
    }
    return hipGetLastError();
}

hipError_t launch_cfl_mixed(const Params &p, const Metrics &m, const Background &bg, const float *h,
                            const unsigned *rr, DevScalars *sc, hipStream_t s) {
    const long long n = (long long)p.npz * p.plane;
    long long nb = (n + 255) / 256;
    if (nb > 4096) nb = 4096;
    if (p.lmag) hipLaunchKernelGGL(k_cfl_m<true>, dim3((unsigned)nb), dim3(256), 0, s, p, m, bg, h, rr, sc);
    else hipLaunchKernelGGL(k_cfl_m<false>, dim3((unsigned)nb), dim3(256), 0, s, p, m, bg, h, rr, sc);
    return hipGetLastError();
}

hipError_t launch_pack_mixed(const Params &p, const double *padded, int field, const Background &bg, float *hreg,
                             unsigned *rreg, hipStream_t s) {
    long long nb = (p.plane * (p.npz + 2) + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_pack_m, dim3((unsigned)nb), dim3(256), 0, s, p, padded, field, bg, hreg, rreg);
    return hipGetLastError();
}

hipError_t launch_unpack_mixed(const Params &p, const float *hreg, const unsigned *rreg, int field,
                               const Background &bg, double *padded, hipStream_t s) {
    const long long n = (long long)(p.npx + 2) * (p.npy + 2) * (p.npz + 2);
    long long nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_unpack_m, dim3((unsigned)nb), dim3(256), 0, s, p, hreg, rreg, field, bg, padded);
    return hipGetLastError();
}

hipError_t launch_rebase_mixed(const Params &p, const Background &ob, const Background &nb, float *hreg,
                               unsigned *rreg, hipStream_t s) {
    const long long n = (long long)p.npz * p.plane;
    long long nb_ = (n + 255) / 256;
    if (nb_ > 8192) nb_ = 8192;
    hipLaunchKernelGGL(k_rebase_m, dim3((unsigned)nb_), dim3(256), 0, s, p, ob, nb, hreg, rreg);
    return hipGetLastError();
}

template <int TY, bool MAG, int MODE, int TXW>
static hipError_t launch_mb(const BoxDev *boxes, const StageItem *items, int nitems, int grid, int a,
                            hipStream_t s) {
    constexpr int lds = mixed_lds_bytes<TY, MAG, MODE, TXW>();
    static std::atomic<unsigned long long> attr{0};
    {
        hipError_t e =
            lds_limit_once(attr, reinterpret_cast<const void *>(&k_stage_m_batch<TY, MAG, MODE, TXW>), lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_stage_m_batch<TY, MAG, MODE, TXW>), dim3(grid), dim3(TXW * TY), lds, s, boxes, items,
                       nitems, a);
    return hipGetLastError();
}

template <int TY, int TXW>
static hipError_t launch_mb_ty(int mode, int lmag, const BoxDev *boxes, const StageItem *items, int nitems,
                               int grid, int a, hipStream_t s) {
    if (lmag) {
        if (mode == 1) return launch_mb<TY, true, 1, TXW>(boxes, items, nitems, grid, a, s);
        if (mode == 2) return launch_mb<TY, true, 2, TXW>(boxes, items, nitems, grid, a, s);
        return launch_mb<TY, true, 3, TXW>(boxes, items, nitems, grid, a, s);
    }
    if (mode == 1) return launch_mb<TY, false, 1, TXW>(boxes, items, nitems, grid, a, s);
    if (mode == 2) return launch_mb<TY, false, 2, TXW>(boxes, items, nitems, grid, a, s);
    return launch_mb<TY, false, 3, TXW>(boxes, items, nitems, grid, a, s);
}

// Tile shapes built for the batched launches: 64x4, 64x8 and 32x8 (the
// caller validates the shape).
hipError_t launch_stage_batch_mixed(int mode, int tile_x, int tile_y, int lmag, const BoxDev *boxes,
                                    const StageItem *items, int nitems, int grid, int a, hipStream_t s) {
    if (tile_x == 32) return launch_mb_ty<8, 32>(mode, lmag, boxes, items, nitems, grid, a, s);
    if (tile_y == 8) return launch_mb_ty<8, MTX>(mode, lmag, boxes, items, nitems, grid, a, s);
    return launch_mb_ty<4, MTX>(mode, lmag, boxes, items, nitems, grid, a, s);
}

hipError_t launch_bcon_batch_mixed(int lmag, const BoxDev *boxes, const ColumnItem *items, int nitems, int a,
                                   bool last, hipStream_t s) {
    dim3 grid(nitems), block(256);
    if (lmag) {
        if (last) hipLaunchKernelGGL((k_bcon_m_batch<true, true>), grid, block, 0, s, boxes, items, a);
        else hipLaunchKernelGGL((k_bcon_m_batch<true, false>), grid, block, 0, s, boxes, items, a);
    } else {
        if (last) hipLaunchKernelGGL((k_bcon_m_batch<false, true>), grid, block, 0, s, boxes, items, a);
        else hipLaunchKernelGGL((k_bcon_m_batch<false, false>), grid, block, 0, s, boxes, items, a);
    }
    return hipGetLastError();
}

hipError_t launch_closure_inputs(const Params &p, const Background &bg, const float *h, const unsigned *rr,
                                 const float *nrm, __bf16 *x, hipStream_t s) {
    long long nb = ((long long)p.npz * p.plane + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_closure_inputs, dim3((unsigned)nb), dim3(256), 0, s, p, bg, h, rr, nrm, x);
    return hipGetLastError();
}

hipError_t launch_closure_forcing(const Params &p, const float *y, const float *nrm_out, float *force, int *force_on,
                                  hipStream_t s) {
    hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(1), 0, s, force_on);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    long long nb = ((long long)p.npz * p.plane + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_closure_forcing, dim3((unsigned)nb), dim3(256), 0, s, p, y, nrm_out, force, force_on);
    return hipGetLastError();
}

}  // namespace rmhd
