// C-ABI of the MI355X compressible-MHD step (declared in include/rast3dmhd.h).
//
// Owns the device state of one resident box.  FP64 form: two q registers
// (ping-pong, because a stage reads the old state's neighbours while writing
// the new one) and the Z register of Wray's low-storage RK3.  Mixed form: two
// FP32 perturbation registers h (ping-pong) and the FP32 register D, plus the
// FP64 background.  Both: the metrics and a small block of per-step scalars.
// One HIP stream per handle.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/rast3dmhd.h"
#include "../../include/rast3dmhd_closure.h"
#include "mhd_host.h"

using namespace rmhd;

struct mhd_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    Params p{};
    mhd_desc desc{};
    int mixed = 0;
    double *q[2] = {nullptr, nullptr};  // FP64: ping-pong state registers
    double *z = nullptr;                // FP64: RK3 Z register
    float *hq[2] = {nullptr, nullptr};  // mixed: FP32 perturbation registers
    float *dr = nullptr;                // mixed: FP32 increment/residual register (within a step)
    unsigned *rr = nullptr;             // mixed: BF16 residual pairs (between steps)
    double *bgd = nullptr;              // mixed: FP64 background arrays
    float *bgf = nullptr;               // mixed: FP32 background arrays
    PlaneM *bgp = nullptr;              // mixed: per-plane records of the stage kernel
    Background bg{};
    std::vector<double> bg_ro, bg_tt;   // mixed: the background profiles as set (npz + 2 planes)
    double *bgd2 = nullptr;             // mixed: second set of background tables (mhd_rebase)
    float *bgf2 = nullptr;
    PlaneM *bgp2 = nullptr;
    double *prof = nullptr;             // mixed: plane means for mhd_rebase [3][npz]
    int rebase_every = 0, since_rebase = 0;
    double *f = nullptr;                // right-hand side (mhd_fluxes only)
    double *stage = nullptr;            // one padded field for host transfers
    double *met = nullptr;              // 11 metric arrays
    Metrics m{};
    std::vector<double> dzzdz, d2zzdz2, ddz, rkapa, dkapa;  // host copies for the background
    DevScalars *sc = nullptr;
    // closure-corrected step (mixed form)
    mhd_net_t *net = nullptr;
    float *force = nullptr;    // [7][fsize]
    int *force_on = nullptr;
    __bf16 *cx = nullptr;      // network input [npz][npy][npx][8]
    float *nrm = nullptr;      // input plane means [8][npz], 1/std[8]; then nrm_out:
    float *nrm_out = nullptr;  // subgrid plane means [7][npz], std[7]
    bool closure = false;
    int cur = 0;
    bool have_metrics = false, have_state = false, have_bg = false;
    bool cfl_valid = false;             // sc->dt/umach describe the current register
    // optional FLUXES/BCON branches (FP64): per-plane tables, relaxation state,
    // the T right-hand side carried between evaluations (LSHR)
    double *optmem = nullptr;
    RelaxState *rs = nullptr;
    double *ftacc = nullptr;
    OptBuffers ob{};
    bool need_coords = false, have_coords = false;
    bool means = false;                 // plane means before every evaluation
    bool timed = false;                 // EXP/TANH of TIMT per step (TC != 0)
    mhd_desc d{};
    char err[512] = {0};
};

static char g_create_err[512];

static int ready(mhd_t *h, const char *who);

static int fail(mhd_t *h, int code, const char *fmt, ...) {
    char *buf = h ? h->err : g_create_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, 512, fmt, ap);
    va_end(ap);
    return code;
}

int rmhd::global_fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_create_err, sizeof g_create_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(h, x)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail((h), MHD_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_),     \
                        __FILE__, __LINE__);                                               \
    } while (0)

static int nfields(const mhd_t *h) { return h->p.lmag ? 8 : 5; }

static void free_all(mhd_t *h) {
    if (!h) return;
    (void)hipSetDevice(h->dev);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void *bufs[25] = {h->q[0], h->q[1], h->z, h->hq[0], h->hq[1], h->dr, h->rr, h->bgd, h->bgf, h->bgp,
                      h->f, h->stage, h->met, h->sc, h->force, h->force_on, h->cx, h->nrm,
                      h->optmem, h->rs, h->ftacc, h->bgd2, h->bgf2, h->bgp2, h->prof};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

// Device buffers of the optional branches: the per-plane and per-axis tables
// (ROM, WWZ, TTM, HRAD, radd, the Gaussians GX, GY, GZ, ZEE, the CFL
// conductivity), the slab sums and relaxation work space, the relaxation
// state (DATA values, sub:1241-1244) and, with LSHR, the carried FT.
static hipError_t create_branches(mhd_t *h) {
    Params &p = h->p;
    const size_t npl = (size_t)p.npz + 2;
    const size_t n = 8 * npl + (p.npx + 2) + (p.npy + 2) + opt_part_len(p) + opt_work_len(p);
    hipError_t e = hipMalloc(&h->optmem, n * sizeof(double));
    if (e == hipSuccess) e = hipMemset(h->optmem, 0, n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&h->rs, sizeof(RelaxState));
    if (e == hipSuccess) {
        const RelaxState r0{1.0, -1e99, 1e99, 0};
        e = hipMemcpy(h->rs, &r0, sizeof(r0), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && p.lshr) {
        e = hipMalloc(&h->ftacc, (size_t)p.fsize * sizeof(double));
        if (e == hipSuccess) e = hipMemset(h->ftacc, 0, (size_t)p.fsize * sizeof(double));  // FT starts at zero
    }
    if (e != hipSuccess) return e;
    double *b = h->optmem;
    OptBuffers &o = h->ob;
    o.rom = b; o.wwz = b + npl; o.ttm = b + 2 * npl; o.hrad = b + 3 * npl; o.radd = b + 4 * npl;
    double *gz = b + 5 * npl, *zee = b + 6 * npl;
    double *gx = b + 8 * npl, *gy = gx + (p.npx + 2);
    o.part = gy + (p.npy + 2);
    o.work = o.part + opt_part_len(p);
    o.rs = h->rs;
    o.zee = zee;
    p.rom = o.rom; p.wwz = o.wwz; p.ttm = o.ttm; p.hrad = o.hrad; p.radd = o.radd;
    p.gx = gx; p.gy = gy; p.gz = gz;
    p.ftacc = h->ftacc;
    h->need_coords = desc_needs_coords(&h->d);
    h->means = p.id != 0 || p.lrem || p.rlax > 0.0;
    h->timed = p.tc != 0.0 && ((p.itc == 2) || (p.tp != 0.0 && (p.itc == 0 || p.itc == 1)));
    return hipSuccess;
}

extern "C" int mhd_create(const mhd_desc *d, mhd_t **out) {
    if (!d || !out) return fail(nullptr, MHD_ERR_ARG, "mhd_create: null argument");
    *out = nullptr;
    {
        char msg[400];
        const int rc = check_desc(d, msg, sizeof msg);
        if (rc) return fail(nullptr, rc, "mhd_create: %s", msg);
    }
    mhd_t *h = new (std::nothrow) mhd_ctx();
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_create: out of host memory");
    h->desc = *d;
    h->d = *d;
    h->mixed = d->precision == MHD_MIXED;
    h->p = params_from_desc(d);
    const Params &p = h->p;
    int rc;
    if (hipGetDevice(&h->dev) != hipSuccess) {
        rc = fail(nullptr, MHD_ERR_HIP, "mhd_create: no HIP device");
        delete h;
        return rc;
    }
    const int nf = p.lmag ? 8 : 5;
    const size_t pbytes = (size_t)(d->npx + 2) * (d->npy + 2) * (d->npz + 2) * sizeof(double);
    const size_t nzp = (size_t)d->npz + 2;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (!h->mixed) {
        const size_t fbytes = (size_t)p.fsize * sizeof(double) * nf;
        if (e == hipSuccess) e = hipMalloc(&h->q[0], fbytes);
        if (e == hipSuccess) e = hipMalloc(&h->q[1], fbytes);
        if (e == hipSuccess) e = hipMalloc(&h->z, fbytes);
    } else {
        const size_t fbytes = (size_t)p.fsize * sizeof(float) * nf;
        if (e == hipSuccess) e = hipMalloc(&h->hq[0], fbytes);
        if (e == hipSuccess) e = hipMalloc(&h->hq[1], fbytes);
        if (e == hipSuccess) e = hipMalloc(&h->dr, fbytes);
        if (e == hipSuccess) e = hipMalloc(&h->rr, resid_bytes(nf, p.fsize));
        if (e == hipSuccess) e = hipMalloc(&h->bgd, BG_DBL * nzp * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&h->bgf, BG_FLT * nzp * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&h->bgp, nzp * sizeof(PlaneM));
    }
    if (e == hipSuccess) e = hipMalloc(&h->stage, pbytes);
    if (e == hipSuccess) e = hipMalloc(&h->met, sizeof(double) * (3 * (d->npx + 2) + 3 * (d->npy + 2) + 5 * (d->npz + 2)));
    if (e == hipSuccess) e = hipMalloc(&h->sc, sizeof(DevScalars));
    if (e != hipSuccess) {
        rc = fail(nullptr, MHD_ERR_HIP, "mhd_create: device allocation failed: %s", hipGetErrorString(e));
        free_all(h);
        delete h;
        return rc;
    }
    if (e == hipSuccess && desc_branches(d)) e = create_branches(h);
    if (e != hipSuccess) {
        rc = fail(nullptr, MHD_ERR_HIP, "mhd_create: device allocation failed: %s", hipGetErrorString(e));
        free_all(h);
        delete h;
        return rc;
    }
    const DevScalars s0 = initial_scalars();
    e = hipMemcpy(h->sc, &s0, sizeof(s0), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        rc = fail(nullptr, MHD_ERR_HIP, "mhd_create: %s", hipGetErrorString(e));
        free_all(h);
        delete h;
        return rc;
    }
    *out = h;
    return MHD_OK;
}

extern "C" int mhd_set_metrics(mhd_t *h, const double *dxxdx, const double *d2xxdx2, const double *ddx,
                               const double *dyydy, const double *d2yydy2, const double *ddy,
                               const double *dzzdz, const double *d2zzdz2, const double *ddz,
                               const double *rkapa, const double *dkapa) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_set_metrics: null handle");
    const double *src[11] = {dxxdx, d2xxdx2, ddx, dyydy, d2yydy2, ddy, dzzdz, d2zzdz2, ddz, rkapa, dkapa};
    size_t len[11];
    const Metrics mt = metric_table(h->p, h->met, len);
    HIPCHK(h, hipSetDevice(h->dev));
    size_t off = 0;
    for (int n = 0; n < 11; ++n) {
        if (!src[n]) return fail(h, MHD_ERR_ARG, "mhd_set_metrics: array %d is null", n);
        HIPCHK(h, hipMemcpyAsync(h->met + off, src[n], len[n] * sizeof(double), hipMemcpyHostToDevice, h->stream));
        off += len[n];
    }
    h->m = mt;
    if (h->p.lrem) {
        // CFL conductivity 1+RKAPA (sub:844-852), in the slot after ZEE
        std::vector<double> rk1((size_t)h->p.npz + 2);
        for (size_t k = 0; k < rk1.size(); ++k) rk1[k] = 1.0 + rkapa[k];
        double *dst = h->optmem + 7 * ((size_t)h->p.npz + 2);
        HIPCHK(h, hipMemcpyAsync(dst, rk1.data(), rk1.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        h->m.rkcfl = dst;
    }
    // the grid-only CFL term, which the stage kernels take from here
    const double d2 = dd2_min(h->p, ddx, ddy, ddz);
    HIPCHK(h, hipMemcpyAsync(reinterpret_cast<char *>(h->sc) + offsetof(DevScalars, dd2min), &d2, sizeof(d2),
                             hipMemcpyHostToDevice, h->stream));
    const size_t nz = (size_t)h->p.npz + 2;
    h->dzzdz.assign(dzzdz, dzzdz + nz);
    h->d2zzdz2.assign(d2zzdz2, d2zzdz2 + nz);
    h->ddz.assign(ddz, ddz + nz);
    h->rkapa.assign(rkapa, rkapa + nz);
    h->dkapa.assign(dkapa, dkapa + nz);
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->have_metrics = true;
    return MHD_OK;
}

extern "C" int mhd_set_coordinates(mhd_t *h, const double *exx, const double *wyy, const double *zee) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_set_coordinates: null handle");
    if (!exx || !wyy || !zee) return fail(h, MHD_ERR_ARG, "mhd_set_coordinates: null array");
    if (!h->optmem) return MHD_OK;  // no branch reads them
    const Params &p = h->p;
    const mhd_desc &d = h->d;
    const size_t nx = p.npx + 2, ny = p.npy + 2, nz = (size_t)p.npz + 2;
    // CLN = -4.0E00*LOG(2.0E00)/HH/HH: the float32 product, then double
    // divisions (sub:1785, 2803); EXP(CLN*(EXX(I)-XP)**2) per index (the host C
    // library's exp, as the reference's Fortran runtime)
    const double cln = -((double)(4.0f * logf(2.0f)) / d.hh / d.hh);
    std::vector<double> gx(nx), gy(ny), gz(nz), wwz(nz);
    for (size_t i = 0; i < nx; ++i) gx[i] = exp(cln * ((exx[i] - d.xp) * (exx[i] - d.xp)));
    for (size_t j = 0; j < ny; ++j) gy[j] = exp(cln * ((wyy[j] - d.yp) * (wyy[j] - d.yp)));
    for (size_t k = 0; k < nz; ++k) gz[k] = exp(cln * ((zee[k] - d.zp) * (zee[k] - d.zp)));
    HIPCHK(h, hipSetDevice(h->dev));
    if (d.id == 2) {
        // WWZ = EXP(CLN*ZEE**2)+EXP(CLN*(ZEE-ZMAX)**2), SGM=0.1E00 (sub:1327-1331)
        const double sgm = (double)0.1f;
        const double c2 = -((double)(4.0f * logf(2.0f)) / sgm / sgm);
        for (size_t k = 0; k < nz; ++k) {
            const double zm = zee[k] - d.zmax;
            wwz[k] = exp(c2 * (zee[k] * zee[k])) + exp(c2 * (zm * zm));
        }
        HIPCHK(h, hipMemcpyAsync(h->ob.wwz, wwz.data(), nz * sizeof(double), hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(h, hipMemcpyAsync(const_cast<double *>(p.gx), gx.data(), nx * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(const_cast<double *>(p.gy), gy.data(), ny * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(const_cast<double *>(p.gz), gz.data(), nz * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(const_cast<double *>(h->ob.zee), zee, nz * sizeof(double), hipMemcpyHostToDevice,
                             h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->have_coords = true;
    return MHD_OK;
}

// The factors of the time-dependent branches at this step's TIMT (the
// reference evaluates them with its Fortran runtime's TANH and EXP; the host
// C library here): the heat-loss ramp 1+TANH(LOG(3.0E00)/QFH*(TIMT-TC))
// (sub:1791) and the plume TPR = TP+(1-TP)*EXP(-TIMT/TC) (sub:2794-2796).
static hipError_t time_factors(mhd_t *h) {
    if (!h->timed) return hipSuccess;
    double timt = 0.0;
    hipError_t e = hipMemcpyAsync(&timt, reinterpret_cast<const char *>(h->sc) + offsetof(DevScalars, timt),
                                  sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return e;
    const mhd_desc &d = h->d;
    h->p.ramp = 1.0 + tanh((double)logf(3.0f) / d.qfh * (timt - d.tc));
    h->p.tpr = d.tp + (1.0 - d.tp) * exp(-(timt / d.tc));
    return hipSuccess;
}

extern "C" int mhd_set_background(mhd_t *h, const double *ro0, const double *tt0) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_set_background: null handle");
    if (!h->mixed) return fail(h, MHD_ERR_STATE, "mhd_set_background: only for MHD_MIXED handles");
    if (!h->have_metrics) return fail(h, MHD_ERR_STATE, "mhd_set_background: set the metrics first");
    if (!ro0 || !tt0) return fail(h, MHD_ERR_ARG, "mhd_set_background: null array");
    std::vector<double> d;
    std::vector<float> f;
    std::vector<PlaneM> pl;
    const int bad = background_tables(h->p, h->dzzdz.data(), h->d2zzdz2.data(), h->ddz.data(), h->rkapa.data(),
                                      h->dkapa.data(), ro0, tt0, d, f, pl);
    if (bad >= 0) return fail(h, MHD_ERR_ARG, "mhd_set_background: rho0[%d] <= 0", bad);
    HIPCHK(h, hipSetDevice(h->dev));
    HIPCHK(h, hipMemcpyAsync(h->bgd, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->bgf, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->bgp, pl.data(), pl.size() * sizeof(PlaneM), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->bg = background_view(h->p.npz, h->bgd, h->bgf);
    h->bg.planes = h->bgp;
    if (h->closure) {
        h->bg.force = h->force;
        h->bg.force_on = h->force_on;
    }
    h->bg_ro.assign(ro0, ro0 + h->p.npz + 2);
    h->bg_tt.assign(tt0, tt0 + h->p.npz + 2);
    h->since_rebase = 0;
    h->have_bg = true;
    h->have_state = false;
    return MHD_OK;
}

// Re-centres the mixed form's background on the current horizontal means of
// rho and T (interior planes; the z ghost planes keep theirs) and re-expresses
// the stored rho and T about it in place, as download + set_background +
// upload would, without the host round trip of the state.
static int rebase(mhd_t *h) {
    const Params &p = h->p;
    const int nz = p.npz;
    const size_t nzp = (size_t)nz + 2;
    if (nz < 3) return fail(h, MHD_ERR_ARG, "mhd_rebase: npz=%d < 3", nz);
    {
        hipError_t e = hipSuccess;
        if (!h->bgd2) e = hipMalloc(&h->bgd2, BG_DBL * nzp * sizeof(double));
        if (e == hipSuccess && !h->bgf2) e = hipMalloc(&h->bgf2, BG_FLT * nzp * sizeof(float));
        if (e == hipSuccess && !h->bgp2) e = hipMalloc(&h->bgp2, nzp * sizeof(PlaneM));
        if (e == hipSuccess && !h->prof) e = hipMalloc(&h->prof, 3 * (size_t)nz * sizeof(double));
        if (e != hipSuccess) return fail(h, MHD_ERR_HIP, "mhd_rebase: allocation: %s", hipGetErrorString(e));
    }
    std::vector<double> prof(3 * (size_t)nz);
    HIPCHK(h, launch_flux_profiles_mixed(p, h->hq[h->cur], h->rr, h->bg, h->prof, h->stream));
    HIPCHK(h, hipMemcpyAsync(prof.data(), h->prof, prof.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::vector<double> ro0 = h->bg_ro, tt0 = h->bg_tt;
    for (int k = 1; k <= nz; ++k) {
        tt0[k] = prof[k - 1];
        ro0[k] = prof[nz + k - 1];
    }
    std::vector<double> d;
    std::vector<float> f;
    std::vector<PlaneM> pl;
    const int bad = background_tables(p, h->dzzdz.data(), h->d2zzdz2.data(), h->ddz.data(), h->rkapa.data(),
                                      h->dkapa.data(), ro0.data(), tt0.data(), d, f, pl);
    if (bad >= 0) return fail(h, MHD_ERR_NONFINITE, "mhd_rebase: horizontal mean of rho[%d] <= 0", bad);
    HIPCHK(h, hipMemcpyAsync(h->bgd2, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->bgf2, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->bgp2, pl.data(), pl.size() * sizeof(PlaneM), hipMemcpyHostToDevice, h->stream));
    Background nb = background_view(nz, h->bgd2, h->bgf2);
    nb.planes = h->bgp2;
    nb.force = h->bg.force;
    nb.force_on = h->bg.force_on;
    HIPCHK(h, launch_rebase_mixed(p, h->bg, nb, h->hq[h->cur], h->rr, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::swap(h->bgd, h->bgd2);
    std::swap(h->bgf, h->bgf2);
    std::swap(h->bgp, h->bgp2);
    h->bg = nb;
    h->bg_ro.swap(ro0);
    h->bg_tt.swap(tt0);
    h->cfl_valid = false;  // the next step reduces the CFL limits of the re-expressed state
    h->since_rebase = 0;
    return MHD_OK;
}

extern "C" int mhd_rebase(mhd_t *h) {
    int rc = ready(h, "mhd_rebase");
    if (rc) return rc;
    if (!h->mixed) return fail(h, MHD_ERR_STATE, "mhd_rebase: only for MHD_MIXED handles");
    HIPCHK(h, hipSetDevice(h->dev));
    return rebase(h);
}

extern "C" int mhd_set_rebase_interval(mhd_t *h, int nsteps) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_set_rebase_interval: null handle");
    if (nsteps < 0) return fail(h, MHD_ERR_ARG, "mhd_set_rebase_interval: nsteps=%d < 0", nsteps);
    if (nsteps && !h->mixed) return fail(h, MHD_ERR_STATE, "mhd_set_rebase_interval: only for MHD_MIXED handles");
    if (nsteps && h->p.npz < 3) return fail(h, MHD_ERR_ARG, "mhd_set_rebase_interval: npz=%d < 3", h->p.npz);
    h->rebase_every = nsteps;
    h->since_rebase = 0;
    return MHD_OK;
}

extern "C" int mhd_upload(mhd_t *h, const double *ru, const double *rv, const double *rw, const double *ro,
                          const double *tt, const double *bx, const double *by, const double *bz) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_upload: null handle");
    if (h->mixed && !h->have_bg) return fail(h, MHD_ERR_STATE, "mhd_upload: set the background first");
    const double *src[8] = {ru, rv, rw, ro, tt, bx, by, bz};
    const int nf = nfields(h);
    const size_t pbytes = (size_t)(h->p.npx + 2) * (h->p.npy + 2) * (h->p.npz + 2) * sizeof(double);
    HIPCHK(h, hipSetDevice(h->dev));
    for (int f = 0; f < nf; ++f) {
        if (!src[f]) return fail(h, MHD_ERR_ARG, "mhd_upload: field %d is null", f);
        HIPCHK(h, hipMemcpyAsync(h->stage, src[f], pbytes, hipMemcpyHostToDevice, h->stream));
        const size_t off = (size_t)f * h->p.fsize;
        // both registers carry the z ghost planes, which the stages never write
        if (!h->mixed) {
            HIPCHK(h, launch_pack(h->p, h->stage, h->q[0] + off, h->stream));
            HIPCHK(h, launch_pack(h->p, h->stage, h->q[1] + off, h->stream));
        } else {
            HIPCHK(h, launch_pack_mixed(h->p, h->stage, f, h->bg, h->hq[1], h->rr, h->stream));
            HIPCHK(h, launch_pack_mixed(h->p, h->stage, f, h->bg, h->hq[0], h->rr, h->stream));
        }
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->cur = 0;
    h->have_state = true;
    h->cfl_valid = false;
    return MHD_OK;
}

extern "C" int mhd_download(mhd_t *h, double *ru, double *rv, double *rw, double *ro, double *tt,
                            double *bx, double *by, double *bz) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_download: null handle");
    if (!h->have_state) return fail(h, MHD_ERR_STATE, "mhd_download: no state uploaded");
    double *dst[8] = {ru, rv, rw, ro, tt, bx, by, bz};
    const int nf = nfields(h);
    const size_t pbytes = (size_t)(h->p.npx + 2) * (h->p.npy + 2) * (h->p.npz + 2) * sizeof(double);
    HIPCHK(h, hipSetDevice(h->dev));
    for (int f = 0; f < nf; ++f) {
        if (!dst[f]) continue;
        const size_t off = (size_t)f * h->p.fsize;
        if (!h->mixed) HIPCHK(h, launch_unpack(h->p, h->q[h->cur] + off, h->stage, 1, h->stream));
        else HIPCHK(h, launch_unpack_mixed(h->p, h->hq[h->cur], h->rr, f, h->bg, h->stage, h->stream));
        HIPCHK(h, hipMemcpyAsync(dst[f], h->stage, pbytes, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return MHD_OK;
}

// The launches of one STEP, optionally bracketed by events (ev[0..13]:
// CFL, then stage/BCON pairs of the three stages).  The CFL limits of the
// state a step starts from are reduced by the previous step's stage-3 and
// boundary kernels; only the first step after an upload runs the
// stand-alone reduction.
static hipError_t launches(mhd_t *h, hipEvent_t *ev) {
    const Params &p = h->p;
    hipError_t e = time_factors(h);
    auto rec = [&](int i) {
        if (ev && e == hipSuccess) e = hipEventRecord(ev[i], h->stream);
    };
    rec(0);
    if (!h->cfl_valid) {
        if (!h->mixed) e = launch_cfl(p, h->m, h->q[h->cur], h->sc, h->stream);
        else e = launch_cfl_mixed(p, h->m, h->bg, h->hq[h->cur], h->rr, h->sc, h->stream);
        if (e == hipSuccess) e = launch_cfl_final(p, h->sc, false, h->stream);
    }
    if (e == hipSuccess && h->closure) {
        // subgrid forcing from the network on the state at the start of the step
        const float *y = nullptr;
        e = launch_closure_inputs(p, h->bg, h->hq[h->cur], h->rr, h->nrm, h->cx, h->stream);
        if (e == hipSuccess) e = net_forward_device(h->net, h->stream, p.npx, p.npy, p.npz, h->cx, &y);
        if (e == hipSuccess) e = launch_closure_forcing(p, y, h->nrm_out, h->force, h->force_on, h->stream);
    }
    rec(1);
    int a = h->cur, b = 1 - h->cur;
    for (int st = 1; st <= 3 && e == hipSuccess; ++st) {
        rec(2 + 4 * (st - 1));
        if (h->means) e = launch_opt_means(p, h->m, h->q[a], h->ob, h->sc, h->stream);
        if (e != hipSuccess) break;
        if (!h->mixed) e = launch_stage(st, p, h->m, h->q[a], h->q[b], h->z, nullptr, h->sc, h->stream);
        else e = launch_stage_mixed(st, p, h->m, h->bg, h->hq[a], h->hq[b], h->dr, h->rr, h->sc, h->stream);
        rec(3 + 4 * (st - 1));
        rec(4 + 4 * (st - 1));
        if (e == hipSuccess) {
            if (!h->mixed) e = launch_bcon(p, h->m, h->q[a], h->q[b], h->sc, st == 3, h->stream);
            else e = launch_bcon_mixed(p, h->m, h->bg, h->hq[b], h->dr, h->rr, h->sc, st == 3, h->stream);
        }
        rec(5 + 4 * (st - 1));
        const int t = a;
        a = b;
        b = t;
    }
    if (e == hipSuccess) e = launch_cfl_final(p, h->sc, true, h->stream);
    if (e == hipSuccess) {
        h->cur = 1 - h->cur;
        h->cfl_valid = true;
    }
    return e;
}

static int finish(mhd_t *h, double *dt, double *umach) {
    DevScalars s;
    HIPCHK(h, hipMemcpyAsync(&s, h->sc, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (dt) *dt = s.dt_used;
    if (umach) *umach = s.umach_used;
    if (s.nonfinite)
        return fail(h, MHD_ERR_NONFINITE, "non-finite CFL limit or Mach number (DT=%g, UMACH=%g)", s.dt_used,
                    s.umach_used);
    return MHD_OK;
}

static int ready(mhd_t *h, const char *who) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "%s: null handle", who);
    if (!h->have_metrics) return fail(h, MHD_ERR_STATE, "%s: metrics not set", who);
    if (!h->have_state) return fail(h, MHD_ERR_STATE, "%s: no state uploaded", who);
    if (h->need_coords && !h->have_coords)
        return fail(h, MHD_ERR_STATE, "%s: the branches of this box need mhd_set_coordinates", who);
    return MHD_OK;
}

// One FLUXES evaluation of the FP64 form into h->f (plane means first).
static hipError_t rhs_eval(mhd_t *h) {
    hipError_t e = hipSuccess;
    if (!h->f) e = hipMalloc(&h->f, (size_t)h->p.fsize * sizeof(double) * nfields(h));
    if (e == hipSuccess) e = time_factors(h);
    if (e == hipSuccess && h->means) e = launch_opt_means(h->p, h->m, h->q[h->cur], h->ob, h->sc, h->stream);
    if (e == hipSuccess) e = launch_stage(0, h->p, h->m, h->q[h->cur], nullptr, nullptr, h->f, h->sc, h->stream);
    return e;
}

// After a step: the periodic re-centring of the mixed background, if set.
static int after_step(mhd_t *h) {
    if (!h->rebase_every || ++h->since_rebase < h->rebase_every) return MHD_OK;
    return rebase(h);
}

extern "C" int mhd_step(mhd_t *h, double *dt, double *umach) {
    int rc = ready(h, "mhd_step");
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->dev));
    HIPCHK(h, launches(h, nullptr));
    rc = finish(h, dt, umach);
    return rc ? rc : after_step(h);
}

extern "C" int mhd_steps(mhd_t *h, int n, double *dt, double *umach) {
    int rc = ready(h, "mhd_steps");
    if (rc) return rc;
    if (n < 0) return fail(h, MHD_ERR_ARG, "mhd_steps: n=%d", n);
    HIPCHK(h, hipSetDevice(h->dev));
    for (int i = 0; i < n; ++i) {
        HIPCHK(h, launches(h, nullptr));
        if (h->rebase_every && ++h->since_rebase >= h->rebase_every) {
            // the step's DT and UMACH are read (and checked) before the re-centring
            rc = finish(h, dt, umach);
            if (rc == MHD_OK) rc = rebase(h);
            if (rc) return rc;
        }
    }
    return finish(h, dt, umach);
}

extern "C" int mhd_time_steps(mhd_t *h, int n, double *ms_step, double *ms_stage, double *ms_cfl,
                              double *ms_bcon) {
    int rc = ready(h, "mhd_time_steps");
    if (rc) return rc;
    if (n < 1) return fail(h, MHD_ERR_ARG, "mhd_time_steps: n=%d", n);
    HIPCHK(h, hipSetDevice(h->dev));
    constexpr int PER = 14;
    std::vector<hipEvent_t> ev((size_t)PER * n + 2);
    size_t made = 0;
    hipError_t e = hipSuccess;
    for (; made < ev.size() && e == hipSuccess; ++made) e = hipEventCreate(&ev[made]);
    if (e != hipSuccess) {
        for (size_t i = 0; i + 1 < made; ++i) (void)hipEventDestroy(ev[i]);
        return fail(h, MHD_ERR_HIP, "mhd_time_steps: hipEventCreate: %s", hipGetErrorString(e));
    }
    e = hipEventRecord(ev[0], h->stream);
    for (int s = 0; s < n && e == hipSuccess; ++s) e = launches(h, &ev[2 + (size_t)PER * s]);
    if (e == hipSuccess) e = hipEventRecord(ev[1], h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    double tstage = 0, tcfl = 0, tbcon = 0;
    float ms = 0, total = 0;
    if (e == hipSuccess) {
        for (int s = 0; s < n; ++s) {
            hipEvent_t *x = &ev[2 + (size_t)PER * s];
            (void)hipEventElapsedTime(&ms, x[0], x[1]);
            tcfl += ms;
            for (int st = 0; st < 3; ++st) {
                (void)hipEventElapsedTime(&ms, x[2 + 4 * st], x[3 + 4 * st]);
                tstage += ms;
                (void)hipEventElapsedTime(&ms, x[4 + 4 * st], x[5 + 4 * st]);
                tbcon += ms;
            }
        }
        (void)hipEventElapsedTime(&total, ev[0], ev[1]);
    }
    for (auto &x : ev) (void)hipEventDestroy(x);
    if (e != hipSuccess) return fail(h, MHD_ERR_HIP, "mhd_time_steps: %s", hipGetErrorString(e));
    if (ms_step) *ms_step = total / n;
    if (ms_stage) *ms_stage = tstage / (3.0 * n);
    if (ms_cfl) *ms_cfl = tcfl / n;
    if (ms_bcon) *ms_bcon = tbcon / (3.0 * n);
    return finish(h, nullptr, nullptr);
}

extern "C" int mhd_get_clock(mhd_t *h, int *nit, double *timc, double *timt) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_get_clock: null handle");
    DevScalars s;
    HIPCHK(h, hipSetDevice(h->dev));
    HIPCHK(h, hipMemcpyAsync(&s, h->sc, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (nit) *nit = s.nit;
    if (timc) *timc = s.timc;
    if (timt) *timt = s.timt;
    return MHD_OK;
}

extern "C" int mhd_set_clock(mhd_t *h, int nit, double timc, double timi) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_set_clock: null handle");
    DevScalars s;
    HIPCHK(h, hipSetDevice(h->dev));
    HIPCHK(h, hipMemcpyAsync(&s, h->sc, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    s.nit = nit;
    s.timc = timc;
    s.timi = timi;
    s.timt = timi + timc;
    HIPCHK(h, hipMemcpyAsync(h->sc, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return MHD_OK;
}

extern "C" int mhd_set_fixed_dt(mhd_t *h, double dt) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_set_fixed_dt: null handle");
    if (!(dt >= 0.0) || dt == INFINITY) return fail(h, MHD_ERR_ARG, "mhd_set_fixed_dt: dt=%g", dt);
    DevScalars s;
    HIPCHK(h, hipSetDevice(h->dev));
    HIPCHK(h, hipMemcpyAsync(&s, h->sc, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    s.fixed_dt = dt;
    if (dt > 0.0) s.dt = dt;
    else h->cfl_valid = false;  // the next step reduces the CFL DT again
    HIPCHK(h, hipMemcpyAsync(h->sc, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return MHD_OK;
}

extern "C" int mhd_fluxes(mhd_t *h, double *fu, double *fv, double *fw, double *fr, double *ft,
                          double *ww1, double *ww2, double *ww3) {
    int rc = ready(h, "mhd_fluxes");
    if (rc) return rc;
    if (h->mixed) return fail(h, MHD_ERR_SCOPE, "mhd_fluxes: only for MHD_FP64 handles");
    HIPCHK(h, hipSetDevice(h->dev));
    const int nf = nfields(h);
    HIPCHK(h, rhs_eval(h));
    double *dst[8] = {fu, fv, fw, fr, ft, ww1, ww2, ww3};
    const size_t pbytes = (size_t)(h->p.npx + 2) * (h->p.npy + 2) * (h->p.npz + 2) * sizeof(double);
    for (int f = 0; f < nf; ++f) {
        if (!dst[f]) continue;
        HIPCHK(h, launch_unpack(h->p, h->f + (size_t)f * h->p.fsize, h->stage, 0, h->stream));
        HIPCHK(h, hipMemcpyAsync(dst[f], h->stage, pbytes, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return MHD_OK;
}

extern "C" int mhd_set_closure(mhd_t *h, mhd_net_t *net, const double *in_profile, const double *in_std,
                               const double *out_profile, const double *out_std, double val_error,
                               double threshold) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_set_closure: null handle");
    if (!h->mixed) return fail(h, MHD_ERR_SCOPE, "mhd_set_closure: only for MHD_MIXED handles");
    HIPCHK(h, hipSetDevice(h->dev));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->closure = false;
    h->net = nullptr;
    h->bg.force = nullptr;
    h->bg.force_on = nullptr;
    // fallback: no network, or one whose validation error exceeds the threshold
    if (!net || !(val_error <= threshold)) return MHD_OK;
    if (!in_profile || !in_std || !out_profile || !out_std)
        return fail(h, MHD_ERR_ARG, "mhd_set_closure: null normalisation");
    int cin = 0, cout = 0;
    net_shape(net, &cin, &cout);
    if (cin != 8 || cout != 7)
        return fail(h, MHD_ERR_ARG, "mhd_set_closure: network maps %d -> %d channels (need 8 -> 7)", cin, cout);
    for (int c = 0; c < 15; ++c)
        if (!((c < 8 ? in_std[c] : out_std[c - 8]) > 0.0))
            return fail(h, MHD_ERR_ARG, "mhd_set_closure: standard deviations must be positive");
    const Params &p = h->p;
    const long long nvox = (long long)p.npz * p.plane;
    int ndev = -1;
    if (net_prepare(net, nvox, &ndev) != MHD_OK) return fail(h, MHD_ERR_HIP, "mhd_set_closure: network workspace");
    if (ndev != h->dev) return fail(h, MHD_ERR_ARG, "mhd_set_closure: network on device %d, box on %d", ndev, h->dev);
    const int nz = p.npz;
    const size_t nin = 8 * (size_t)nz + 8, nout = 7 * (size_t)nz + 7;
    if (!h->force) HIPCHK(h, hipMalloc(&h->force, (size_t)7 * p.fsize * sizeof(float)));
    if (!h->force_on) HIPCHK(h, hipMalloc(&h->force_on, sizeof(int)));
    if (!h->cx) HIPCHK(h, hipMalloc(&h->cx, (size_t)nvox * 8 * sizeof(__bf16)));
    if (!h->nrm) HIPCHK(h, hipMalloc(&h->nrm, (nin + nout) * sizeof(float)));
    std::vector<float> nrm(nin + nout);
    for (size_t i = 0; i < 8 * (size_t)nz; ++i) nrm[i] = (float)in_profile[i];
    for (int c = 0; c < 8; ++c) nrm[8 * nz + c] = (float)(1.0 / in_std[c]);
    for (size_t i = 0; i < 7 * (size_t)nz; ++i) nrm[nin + i] = (float)out_profile[i];
    for (int c = 0; c < 7; ++c) nrm[nin + 7 * nz + c] = (float)out_std[c];
    HIPCHK(h, hipMemcpy(h->nrm, nrm.data(), nrm.size() * sizeof(float), hipMemcpyHostToDevice));
    h->nrm_out = h->nrm + nin;
    // forcing is zero outside the interior planes
    HIPCHK(h, hipMemset(h->force, 0, (size_t)7 * p.fsize * sizeof(float)));
    h->net = net;
    h->closure = true;
    if (h->have_bg) {
        h->bg.force = h->force;
        h->bg.force_on = h->force_on;
    }
    return MHD_OK;
}

extern "C" int mhd_closure_active(const mhd_t *h) { return h && h->closure ? 1 : 0; }

extern "C" int mhd_box_filter(mhd_t *h, int factor, int zmode, int source, double *out) {
    int rc = ready(h, "mhd_box_filter");
    if (rc) return rc;
    if (h->mixed) return fail(h, MHD_ERR_SCOPE, "mhd_box_filter: only for MHD_FP64 handles");
    if (!out) return fail(h, MHD_ERR_ARG, "mhd_box_filter: null output");
    const Params &p = h->p;
    if (factor < 1 || p.npx % factor || p.npy % factor || (zmode != 0 && zmode != 1) || (source != 0 && source != 1))
        return fail(h, MHD_ERR_ARG, "mhd_box_filter: factor %d, zmode %d, source %d for %dx%dx%d", factor, zmode,
                    source, p.npx, p.npy, p.npz);
    if (zmode == 0 ? p.npz % factor != 0 : ((p.npz - 1) % factor != 0 || factor % 2 != 0))
        return fail(h, MHD_ERR_ARG, "mhd_box_filter: npz=%d does not coarsen by %d (zmode %d)", p.npz, factor, zmode);
    const int nxc = p.npx / factor, nyc = p.npy / factor;
    const int nzc = zmode == 0 ? p.npz / factor : (p.npz - 1) / factor + 1;
    const int nf = nfields(h);
    HIPCHK(h, hipSetDevice(h->dev));
    const double *src = h->q[h->cur];
    if (source == 1) {
        HIPCHK(h, rhs_eval(h));
        src = h->f;
    }
    const size_t bytes = (size_t)nf * nzc * nyc * nxc * sizeof(double);
    double *d = nullptr;
    HIPCHK(h, hipMalloc(&d, bytes));
    hipError_t e = launch_box_filter(p, src, nf, factor, zmode, nxc, nyc, nzc, d, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(h, MHD_ERR_HIP, "mhd_box_filter: %s", hipGetErrorString(e));
    return MHD_OK;
}

extern "C" int mhd_flux_profiles(mhd_t *h, double *ttm, double *rom, double *fconm, double *fradm) {
    int rc = ready(h, "mhd_flux_profiles");
    if (rc) return rc;
    const Params &p = h->p;
    if (p.npz < 3) return fail(h, MHD_ERR_ARG, "mhd_flux_profiles: npz=%d < 3", p.npz);
    HIPCHK(h, hipSetDevice(h->dev));
    const int nz = p.npz;
    std::vector<double> prof(3 * (size_t)nz);
    double *d = nullptr;
    HIPCHK(h, hipMalloc(&d, prof.size() * sizeof(double)));
    hipError_t e = h->mixed ? launch_flux_profiles_mixed(p, h->hq[h->cur], h->rr, h->bg, d, h->stream)
                            : launch_flux_profiles(p, h->q[h->cur], d, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(prof.data(), d, prof.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(h, MHD_ERR_HIP, "mhd_flux_profiles: %s", hipGetErrorString(e));
    const double *t = prof.data();  // t[k-1]: plane k = 1..nz (reference K = k+1)
    if (ttm) std::copy(t, t + nz, ttm);
    if (rom) std::copy(t + nz, t + 2 * nz, rom);
    if (fconm) std::copy(t + 2 * nz, t + 3 * nz, fconm);
    if (fradm) {
        // FRADM (3dmhdsub.f:1663-1677): central differences, one-sided on
        // the two boundary planes; metric arrays are indexed by padded plane
        const double HZ = p.hz, REPR = p.repr;
        const double *dz = h->dzzdz.data(), *rk = h->rkapa.data();
        for (int k = 2; k <= nz - 1; ++k) fradm[k - 1] = (t[k] - t[k - 2]) * HZ * dz[k] * rk[k] / REPR;
        fradm[0] = (-(3.0 * t[0]) + 4.0 * t[1] - t[2]) * HZ * dz[1] * rk[1] / REPR;
        fradm[nz - 1] = (3.0 * t[nz - 1] - 4.0 * t[nz - 2] + t[nz - 3]) * HZ * dz[nz] * rk[nz] / REPR;
    }
    return MHD_OK;
}

extern "C" const char *mhd_last_error(const mhd_t *h) { return h ? h->err : g_create_err; }

extern "C" void mhd_destroy(mhd_t *h) {
    if (!h) return;
    free_all(h);
    delete h;
}
