// C-ABI of the MI355X compressible-MHD step (declared in include/rast3dmhd.h).
//
// Owns the device state of one resident box: two q registers (ping-pong,
// because a stage reads neighbours of the old state while writing the new
// one), the Z register of Wray's low-storage RK3, the metrics, and a small
// block of per-step scalars.  One HIP stream per handle.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

#include "../../include/rast3dmhd.h"
#include "mhd_kernels.h"

using namespace rmhd;

struct mhd_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    Params p{};
    mhd_desc desc{};
    double *q[2] = {nullptr, nullptr};  // ping-pong state registers
    double *z = nullptr;                // RK3 Z register
    double *f = nullptr;                // right-hand side (mhd_fluxes only)
    double *stage = nullptr;            // one padded field for host transfers
    double *met = nullptr;              // 11 metric arrays
    Metrics m{};
    DevScalars *sc = nullptr;
    int cur = 0;
    bool have_metrics = false, have_state = false;
    bool cfl_valid = false;             // sc->dt/umach describe q[cur]
    char err[512] = {0};
};

static char g_create_err[512];

static int fail(mhd_t *h, int code, const char *fmt, ...) {
    char *buf = h ? h->err : g_create_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, 512, fmt, ap);
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
    void *bufs[7] = {h->q[0], h->q[1], h->z, h->f, h->stage, h->met, h->sc};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

extern "C" int mhd_create(const mhd_desc *d, mhd_t **out) {
    if (!d || !out) return fail(nullptr, MHD_ERR_ARG, "mhd_create: null argument");
    *out = nullptr;
    if (d->npx < 3 || d->npy < 3 || d->npz < 4)
        return fail(nullptr, MHD_ERR_ARG, "mhd_create: box %dx%dx%d too small (need npx,npy>=3, npz>=4)",
                    d->npx, d->npy, d->npz);
    if (d->izcon != 0 && d->izcon != 1)
        return fail(nullptr, MHD_ERR_SCOPE, "mhd_create: IZCON=%d (BCON: Invalid IZCON)", d->izcon);
    if (d->itcon != 0 && d->itcon != 1)
        return fail(nullptr, MHD_ERR_SCOPE, "mhd_create: ITCON=%d not supported (0 or 1)", d->itcon);
    if (d->tp != 0.0)
        return fail(nullptr, MHD_ERR_SCOPE, "mhd_create: plume boundary (TP != 0) not supported");
    mhd_t *h = new (std::nothrow) mhd_ctx();
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_create: out of host memory");
    h->desc = *d;
    Params &p = h->p;
    p.npx = d->npx; p.npy = d->npy; p.npz = d->npz;
    p.lmag = d->lmag ? 1 : 0; p.lrot = d->lrot ? 1 : 0;
    p.izc = d->izcon; p.itc = d->itcon; p.ibc = d->ibcon;
    p.tile_y = 4;
    if (const char *ty = getenv("RAST3DMHD_TILE_Y")) p.tile_y = atoi(ty) == 8 ? 8 : 4;
    p.plane = (long long)d->npx * d->npy;
    p.fsize = p.plane * (d->npz + 2);
    p.hx = d->hx; p.h2x = d->h2x; p.hy = d->hy; p.h2y = d->h2y; p.hz = d->hz; p.h2z = d->h2z;
    p.c13 = d->c13; p.c23 = d->c23; p.c43 = d->c43;
    p.gam1 = d->gam1; p.gam2 = d->gam2; p.gam3 = d->gam3; p.zeta1 = d->zeta1; p.zeta2 = d->zeta2;
    p.cv = d->cv; p.ocv = d->ocv; p.ore = d->ore; p.re = d->re; p.repr = d->repr;
    p.theta = d->theta; p.grav = d->grav; p.sf = d->sf; p.gamma = d->gamma;
    p.omx = d->lrot ? d->omx : 0.0; p.omz = d->lrot ? d->omz : 0.0;
    p.orm = d->orm; p.rm = d->rm; p.obeta = d->obeta; p.tb = d->tb; p.dzf = d->dzf;
    // products formed once per FLUXES / STEP call in the reference
    p.ogamma = 1.0 / d->gamma;             // sub:744
    p.t_cond = d->ocv / d->repr;           // sub:1584
    p.t_visc = d->ocv * d->ore;            // sub:1904
    p.t_visc43 = d->c43 * d->ocv * d->ore; // sub:1962
    p.t_cross = d->c13 * d->ore;           // sub:1930
    p.t_ohm = d->obeta * d->orm * d->ocv;  // sub:2287
    int rc;
    if (hipGetDevice(&h->dev) != hipSuccess) { rc = fail(nullptr, MHD_ERR_HIP, "mhd_create: no HIP device"); delete h; return rc; }
    const size_t fbytes = (size_t)p.fsize * sizeof(double);
    const size_t pbytes = (size_t)(d->npx + 2) * (d->npy + 2) * (d->npz + 2) * sizeof(double);
    const int nf = p.lmag ? 8 : 5;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&h->q[0], fbytes * nf);
    if (e == hipSuccess) e = hipMalloc(&h->q[1], fbytes * nf);
    if (e == hipSuccess) e = hipMalloc(&h->z, fbytes * nf);
    if (e == hipSuccess) e = hipMalloc(&h->stage, pbytes);
    if (e == hipSuccess) e = hipMalloc(&h->met, sizeof(double) * (3 * (d->npx + 2) + 3 * (d->npy + 2) + 5 * (d->npz + 2)));
    if (e == hipSuccess) e = hipMalloc(&h->sc, sizeof(DevScalars));
    if (e != hipSuccess) {
        rc = fail(nullptr, MHD_ERR_HIP, "mhd_create: device allocation failed: %s", hipGetErrorString(e));
        free_all(h);
        delete h;
        return rc;
    }
    DevScalars s0{};
    s0.key_mach = dkey(0.0);
    for (int n = 0; n < 4; ++n) s0.key_min[n] = dkey(INFINITY);
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
    const size_t len[11] = {(size_t)h->p.npx + 2, (size_t)h->p.npx + 2, (size_t)h->p.npx + 2,
                            (size_t)h->p.npy + 2, (size_t)h->p.npy + 2, (size_t)h->p.npy + 2,
                            (size_t)h->p.npz + 2, (size_t)h->p.npz + 2, (size_t)h->p.npz + 2,
                            (size_t)h->p.npz + 2, (size_t)h->p.npz + 2};
    const double **dst[11] = {&h->m.dxxdx, &h->m.d2xxdx2, &h->m.ddx, &h->m.dyydy, &h->m.d2yydy2, &h->m.ddy,
                              &h->m.dzzdz, &h->m.d2zzdz2, &h->m.ddz, &h->m.rkapa, &h->m.dkapa};
    HIPCHK(h, hipSetDevice(h->dev));
    size_t off = 0;
    for (int n = 0; n < 11; ++n) {
        if (!src[n]) return fail(h, MHD_ERR_ARG, "mhd_set_metrics: array %d is null", n);
        HIPCHK(h, hipMemcpyAsync(h->met + off, src[n], len[n] * sizeof(double), hipMemcpyHostToDevice, h->stream));
        *dst[n] = h->met + off;
        off += len[n];
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->have_metrics = true;
    return MHD_OK;
}

extern "C" int mhd_upload(mhd_t *h, const double *ru, const double *rv, const double *rw, const double *ro,
                          const double *tt, const double *bx, const double *by, const double *bz) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "mhd_upload: null handle");
    const double *src[8] = {ru, rv, rw, ro, tt, bx, by, bz};
    const int nf = nfields(h);
    const size_t pbytes = (size_t)(h->p.npx + 2) * (h->p.npy + 2) * (h->p.npz + 2) * sizeof(double);
    HIPCHK(h, hipSetDevice(h->dev));
    for (int f = 0; f < nf; ++f) {
        if (!src[f]) return fail(h, MHD_ERR_ARG, "mhd_upload: field %d is null", f);
        HIPCHK(h, hipMemcpyAsync(h->stage, src[f], pbytes, hipMemcpyHostToDevice, h->stream));
        // both registers carry the z ghost planes, which the stages never write
        HIPCHK(h, launch_pack(h->p, h->stage, h->q[0] + (size_t)f * h->p.fsize, h->stream));
        HIPCHK(h, launch_pack(h->p, h->stage, h->q[1] + (size_t)f * h->p.fsize, h->stream));
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
        HIPCHK(h, launch_unpack(h->p, h->q[h->cur] + (size_t)f * h->p.fsize, h->stage, 1, h->stream));
        HIPCHK(h, hipMemcpyAsync(dst[f], h->stage, pbytes, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return MHD_OK;
}

// One STEP.  The CFL limits of the state a step starts from are reduced by
// the previous step's stage-3 and boundary kernels; only the first step
// after an upload runs the stand-alone reduction.
static int enqueue_step(mhd_t *h) {
    const Params &p = h->p;
    double *qa = h->q[h->cur], *qb = h->q[1 - h->cur];
    if (!h->cfl_valid) {
        HIPCHK(h, launch_cfl(p, h->m, qa, h->sc, h->stream));
        HIPCHK(h, launch_cfl_final(p, h->sc, false, h->stream));
    }
    HIPCHK(h, launch_stage(1, p, h->m, qa, qb, h->z, nullptr, h->sc, h->stream));
    HIPCHK(h, launch_bcon(p, h->m, qa, qb, h->sc, false, h->stream));
    HIPCHK(h, launch_stage(2, p, h->m, qb, qa, h->z, nullptr, h->sc, h->stream));
    HIPCHK(h, launch_bcon(p, h->m, qb, qa, h->sc, false, h->stream));
    HIPCHK(h, launch_stage(3, p, h->m, qa, qb, h->z, nullptr, h->sc, h->stream));
    HIPCHK(h, launch_bcon(p, h->m, qa, qb, h->sc, true, h->stream));
    HIPCHK(h, launch_cfl_final(p, h->sc, true, h->stream));
    h->cur = 1 - h->cur;
    h->cfl_valid = true;
    return MHD_OK;
}

static int finish(mhd_t *h, double *dt, double *umach) {
    DevScalars s;
    HIPCHK(h, hipMemcpyAsync(&s, h->sc, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (dt) *dt = s.dt_used;
    if (umach) *umach = s.umach_used;
    if (s.nonfinite)
        return fail(h, MHD_ERR_NONFINITE, "non-finite CFL limit or Mach number (DT=%g, UMACH=%g)", s.dt, s.umach);
    return MHD_OK;
}

static int ready(mhd_t *h, const char *who) {
    if (!h) return fail(nullptr, MHD_ERR_ARG, "%s: null handle", who);
    if (!h->have_metrics) return fail(h, MHD_ERR_STATE, "%s: metrics not set", who);
    if (!h->have_state) return fail(h, MHD_ERR_STATE, "%s: no state uploaded", who);
    return MHD_OK;
}

extern "C" int mhd_step(mhd_t *h, double *dt, double *umach) {
    int rc = ready(h, "mhd_step");
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->dev));
    if ((rc = enqueue_step(h))) return rc;
    return finish(h, dt, umach);
}

extern "C" int mhd_steps(mhd_t *h, int n, double *dt, double *umach) {
    int rc = ready(h, "mhd_steps");
    if (rc) return rc;
    if (n < 0) return fail(h, MHD_ERR_ARG, "mhd_steps: n=%d", n);
    HIPCHK(h, hipSetDevice(h->dev));
    for (int i = 0; i < n; ++i)
        if ((rc = enqueue_step(h))) return rc;
    return finish(h, dt, umach);
}

extern "C" int mhd_time_steps(mhd_t *h, int n, double *ms_step, double *ms_stage, double *ms_cfl,
                              double *ms_bcon) {
    int rc = ready(h, "mhd_time_steps");
    if (rc) return rc;
    if (n < 1) return fail(h, MHD_ERR_ARG, "mhd_time_steps: n=%d", n);
    HIPCHK(h, hipSetDevice(h->dev));
    // event pairs: per step, CFL, 3 stages, 3 BCON launches
    const int per = 7;
    hipEvent_t *ev = new (std::nothrow) hipEvent_t[(size_t)2 * per * n + 2];
    if (!ev) return fail(h, MHD_ERR_ARG, "mhd_time_steps: out of host memory");
    const int nev = 2 * per * n + 2;
    int made = 0;
    hipError_t e = hipSuccess;
    for (; made < nev && e == hipSuccess; ++made) e = hipEventCreate(&ev[made]);
    if (e != hipSuccess) {
        for (int i = 0; i < made - 1; ++i) hipEventDestroy(ev[i]);
        delete[] ev;
        return fail(h, MHD_ERR_HIP, "mhd_time_steps: hipEventCreate: %s", hipGetErrorString(e));
    }
    const Params &p = h->p;
    int ie = 2;
    e = hipEventRecord(ev[0], h->stream);
    for (int s = 0; s < n && e == hipSuccess; ++s) {
        double *qa = h->q[h->cur], *qb = h->q[1 - h->cur];
        double *in[3] = {qa, qb, qa}, *out[3] = {qb, qa, qb};
        (void)hipEventRecord(ev[ie++], h->stream);
        if (!h->cfl_valid) {
            e = launch_cfl(p, h->m, qa, h->sc, h->stream);
            if (e == hipSuccess) e = launch_cfl_final(p, h->sc, false, h->stream);
        }
        (void)hipEventRecord(ev[ie++], h->stream);
        for (int st = 0; st < 3 && e == hipSuccess; ++st) {
            (void)hipEventRecord(ev[ie++], h->stream);
            e = launch_stage(st + 1, p, h->m, in[st], out[st], h->z, nullptr, h->sc, h->stream);
            (void)hipEventRecord(ev[ie++], h->stream);
            if (e != hipSuccess) break;
            (void)hipEventRecord(ev[ie++], h->stream);
            e = launch_bcon(p, h->m, in[st], out[st], h->sc, st == 2, h->stream);
            (void)hipEventRecord(ev[ie++], h->stream);
        }
        if (e == hipSuccess) e = launch_cfl_final(p, h->sc, true, h->stream);
        h->cur = 1 - h->cur;
        h->cfl_valid = true;
    }
    if (e == hipSuccess) e = hipEventRecord(ev[1], h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    double tstage = 0, tcfl = 0, tbcon = 0;
    float ms = 0;
    if (e == hipSuccess) {
        int j = 2;
        for (int s = 0; s < n; ++s) {
            hipEventElapsedTime(&ms, ev[j], ev[j + 1]); tcfl += ms; j += 2;
            for (int st = 0; st < 3; ++st) {
                hipEventElapsedTime(&ms, ev[j], ev[j + 1]); tstage += ms; j += 2;
                hipEventElapsedTime(&ms, ev[j], ev[j + 1]); tbcon += ms; j += 2;
            }
        }
        hipEventElapsedTime(&ms, ev[0], ev[1]);
    }
    for (int i = 0; i < nev; ++i) hipEventDestroy(ev[i]);
    delete[] ev;
    if (e != hipSuccess) return fail(h, MHD_ERR_HIP, "mhd_time_steps: %s", hipGetErrorString(e));
    if (ms_step) *ms_step = ms / n;
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

extern "C" int mhd_fluxes(mhd_t *h, double *fu, double *fv, double *fw, double *fr, double *ft,
                          double *ww1, double *ww2, double *ww3) {
    int rc = ready(h, "mhd_fluxes");
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->dev));
    const int nf = nfields(h);
    if (!h->f) HIPCHK(h, hipMalloc(&h->f, (size_t)h->p.fsize * sizeof(double) * nf));
    HIPCHK(h, launch_stage(0, h->p, h->m, h->q[h->cur], nullptr, nullptr, h->f, h->sc, h->stream));
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

extern "C" const char *mhd_last_error(const mhd_t *h) { return h ? h->err : g_create_err; }

extern "C" void mhd_destroy(mhd_t *h) {
    if (!h) return;
    free_all(h);
    delete h;
}
