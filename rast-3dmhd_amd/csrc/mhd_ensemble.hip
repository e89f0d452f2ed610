// C-ABI of the batched ensemble (include/rast3dmhd_ensemble.h).
//
// All boxes live in one device pool: per box two state registers (FP64 q
// ping-pong, or FP32 h ping-pong) and the third register (FP64 Z, or FP32
// D), then the metric blocks, the mixed form's background blocks and one
// DevScalars per box.  The device table of BoxDev entries points every
// batched kernel at its box.  A step is: 3 x (persistent stage launch over
// the StageItem table of all boxes, boundary-plane launch over the
// ColumnItem table), then one per-box DT/clock update.  Steps are replayed
// from captured HIP graphs (one graph per register parity for a pair of
// steps and for a single step); the launches carry no host-side state that
// changes between steps (DT lives in device memory), so a replay is the
// same work as direct launches.

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/rast3dmhd_ensemble.h"
#include "mhd_host.h"

using namespace rmhd;

namespace {

struct EnsBox {
    mhd_desc desc{};
    Params p{};
    size_t reg_elems = 0;            // elements of one register (nf * fsize)
    std::vector<double> dzzdz, d2zzdz2, ddz, rkapa, dkapa;
    bool have_metrics = false, have_bg = false, have_state = false;
    bool cfl_valid = false;
};

constexpr size_t ALIGN = 256;
size_t align_up(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

}  // namespace

struct mhd_ens {
    int dev = 0;
    hipStream_t stream = nullptr;
    int nbox = 0, lmag = 0, mixed = 0, cus = 256;
    std::vector<EnsBox> box;
    std::vector<BoxDev> hbox;        // host copy of the device table
    BoxDev *dbox = nullptr;
    char *pool = nullptr;            // state registers of all boxes
    double *met = nullptr;           // metric blocks
    double *bgd = nullptr;           // mixed: FP64 background blocks
    float *bgf = nullptr;            // mixed: FP32 background blocks
    PlaneM *bgp = nullptr;           // mixed: per-plane records of the stage kernel, per box
    DevScalars *sc = nullptr;
    double *stage = nullptr;         // one padded field of the largest box
    size_t stage_bytes = 0;
    // tiling
    int tile_x = 64, tile_y = 4, kdepth = 16, grid = 0;
    StageItem *ditems = nullptr;
    int nitems = 0;
    ColumnItem *dcols = nullptr;
    int ncols = 0;
    // graphs [single=0 / pair=1][register parity]
    hipGraphExec_t gexec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int cur = 0;
    char err[512] = {0};
};

static char g_ens_err[512];

static int efail(mhd_ens_t *e, int code, const char *fmt, ...) {
    char *buf = e ? e->err : g_ens_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, 512, fmt, ap);
    va_end(ap);
    return code;
}

#define EHIP(e, x)                                                                         \
    do {                                                                                   \
        hipError_t r_ = (x);                                                               \
        if (r_ != hipSuccess)                                                              \
            return efail((e), MHD_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(r_),    \
                         __FILE__, __LINE__);                                              \
    } while (0)

static void drop_graphs(mhd_ens_t *e) {
    for (auto &row : e->gexec)
        for (auto &g : row)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

static void free_ens(mhd_ens_t *e) {
    (void)hipSetDevice(e->dev);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    void *bufs[10] = {e->dbox, e->pool, e->met, e->bgd, e->bgf, e->bgp, e->sc, e->stage, e->ditems, e->dcols};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (e->stream) (void)hipStreamDestroy(e->stream);
}

static int check_box(mhd_ens_t *e, int b, const char *who) {
    if (!e) return efail(nullptr, MHD_ERR_ARG, "%s: null handle", who);
    if (b < 0 || b >= e->nbox) return efail(e, MHD_ERR_ARG, "%s: box %d out of range [0,%d)", who, b, e->nbox);
    return MHD_OK;
}

// Work tables for the current tiling: StageItems (tile_x x tile_y tiles)
// box-major (z chunk, row, x tile), ColumnItems 256 columns each.
static int build_tables(mhd_ens_t *e) {
    std::vector<StageItem> items;
    std::vector<ColumnItem> cols;
    for (int b = 0; b < e->nbox; ++b) {
        const Params &p = e->box[b].p;
        const int kd = e->kdepth > 0 ? e->kdepth : p.npz;
        const int ntx = (p.npx + e->tile_x - 1) / e->tile_x, nty = (p.npy + e->tile_y - 1) / e->tile_y;
        for (int k0 = 1; k0 <= p.npz; k0 += kd)
            for (int ty = 0; ty < nty; ++ty)
                for (int tx = 0; tx < ntx; ++tx) {
                    StageItem it;
                    it.box = b;
                    it.i0 = tx * e->tile_x;
                    it.j0 = ty * e->tile_y;
                    it.kbeg = k0;
                    it.kend = k0 + kd - 1 < p.npz ? k0 + kd - 1 : p.npz;
                    items.push_back(it);
                }
        for (long long c = 0; c < p.plane; c += 256) cols.push_back(ColumnItem{b, (int)c});
    }
    drop_graphs(e);
    if (e->ditems) EHIP(e, hipFree(e->ditems));
    if (e->dcols) EHIP(e, hipFree(e->dcols));
    e->ditems = nullptr;
    e->dcols = nullptr;
    EHIP(e, hipMalloc(&e->ditems, items.size() * sizeof(StageItem)));
    EHIP(e, hipMalloc(&e->dcols, cols.size() * sizeof(ColumnItem)));
    EHIP(e, hipMemcpy(e->ditems, items.data(), items.size() * sizeof(StageItem), hipMemcpyHostToDevice));
    EHIP(e, hipMemcpy(e->dcols, cols.data(), cols.size() * sizeof(ColumnItem), hipMemcpyHostToDevice));
    e->nitems = (int)items.size();
    e->ncols = (int)cols.size();
    return MHD_OK;
}

extern "C" int mhd_ens_create(const mhd_desc *descs, int nbox, mhd_ens_t **out) {
    if (!descs || !out || nbox < 1) return efail(nullptr, MHD_ERR_ARG, "mhd_ens_create: bad arguments");
    *out = nullptr;
    for (int b = 0; b < nbox; ++b) {
        char msg[400];
        const int rc = check_desc(&descs[b], msg, sizeof msg);
        if (rc) return efail(nullptr, rc, "mhd_ens_create: box %d: %s", b, msg);
        if (desc_branches(&descs[b]))
            return efail(nullptr, MHD_ERR_SCOPE, "mhd_ens_create: box %d: the ensemble has the main path only", b);
        if ((descs[b].lmag != 0) != (descs[0].lmag != 0) || descs[b].precision != descs[0].precision)
            return efail(nullptr, MHD_ERR_ARG, "mhd_ens_create: box %d: lmag and precision must match box 0", b);
    }
    mhd_ens_t *e = new (std::nothrow) mhd_ens();
    if (!e) return efail(nullptr, MHD_ERR_ARG, "mhd_ens_create: out of host memory");
    e->nbox = nbox;
    e->lmag = descs[0].lmag ? 1 : 0;
    e->mixed = descs[0].precision == MHD_MIXED;
    e->box.resize(nbox);
    e->hbox.resize(nbox);
    const int nf = e->lmag ? 8 : 5;
    const size_t esz = e->mixed ? sizeof(float) : sizeof(double);
    size_t pool = 0, met = 0, nzsum = 0;
    std::vector<size_t> reg_off(nbox), met_off(nbox), bg_off(nbox);
    for (int b = 0; b < nbox; ++b) {
        EnsBox &B = e->box[b];
        B.desc = descs[b];
        B.p = params_from_desc(&descs[b]);
        B.reg_elems = (size_t)nf * B.p.fsize;
        reg_off[b] = pool;
        pool += 3 * align_up(B.reg_elems * esz) + (e->mixed ? align_up(resid_bytes(nf, B.p.fsize)) : 0);
        met_off[b] = met;
        met += align_up(metric_len(B.p) * sizeof(double)) / sizeof(double);
        bg_off[b] = nzsum;
        nzsum += align_up((size_t)(B.p.npz + 2) * sizeof(double)) / sizeof(double);
        const size_t pb = (size_t)(B.p.npx + 2) * (B.p.npy + 2) * (B.p.npz + 2) * sizeof(double);
        if (pb > e->stage_bytes) e->stage_bytes = pb;
    }
    hipError_t r = hipGetDevice(&e->dev);
    if (r == hipSuccess) r = hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, e->dev);
    if (r == hipSuccess) r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (r == hipSuccess) r = hipMalloc(&e->pool, pool);
    if (r == hipSuccess) r = hipMalloc(&e->met, met * sizeof(double));
    if (r == hipSuccess && e->mixed) r = hipMalloc(&e->bgd, BG_DBL * nzsum * sizeof(double));
    if (r == hipSuccess && e->mixed) r = hipMalloc(&e->bgf, BG_FLT * nzsum * sizeof(float));
    if (r == hipSuccess && e->mixed) r = hipMalloc(&e->bgp, nzsum * sizeof(PlaneM));
    if (r == hipSuccess) r = hipMalloc(&e->sc, nbox * sizeof(DevScalars));
    if (r == hipSuccess) r = hipMalloc(&e->stage, e->stage_bytes);
    if (r == hipSuccess) r = hipMalloc(&e->dbox, nbox * sizeof(BoxDev));
    if (r != hipSuccess) {
        const int rc = efail(nullptr, MHD_ERR_HIP, "mhd_ens_create: %s", hipGetErrorString(r));
        free_ens(e);
        delete e;
        return rc;
    }
    std::vector<DevScalars> s0(nbox, initial_scalars());
    for (int b = 0; b < nbox; ++b) {
        EnsBox &B = e->box[b];
        BoxDev &D = e->hbox[b];
        D.p = B.p;
        size_t len[11];
        D.m = metric_table(B.p, e->met + met_off[b], len);
        char *base = e->pool + reg_off[b];
        const size_t rb = align_up(B.reg_elems * esz);
        if (e->mixed) {
            D.h[0] = reinterpret_cast<float *>(base);
            D.h[1] = reinterpret_cast<float *>(base + rb);
            D.dr = reinterpret_cast<float *>(base + 2 * rb);
            D.rr = reinterpret_cast<unsigned *>(base + 3 * rb);
            const size_t o = bg_off[b];
            // BG_DBL FP64 and BG_FLT FP32 tables of npz+2 entries, laid out per box
            D.bg = background_view(B.p.npz, e->bgd + BG_DBL * o, e->bgf + BG_FLT * o);
            D.bg.planes = e->bgp + o;
        } else {
            D.q[0] = reinterpret_cast<double *>(base);
            D.q[1] = reinterpret_cast<double *>(base + rb);
            D.z = reinterpret_cast<double *>(base + 2 * rb);
        }
        D.sc = e->sc + b;
    }
    r = hipMemcpy(e->sc, s0.data(), nbox * sizeof(DevScalars), hipMemcpyHostToDevice);
    if (r == hipSuccess) r = hipMemcpy(e->dbox, e->hbox.data(), nbox * sizeof(BoxDev), hipMemcpyHostToDevice);
    int rc = r == hipSuccess ? MHD_OK : efail(nullptr, MHD_ERR_HIP, "mhd_ens_create: %s", hipGetErrorString(r));
    if (!rc) {
        rc = build_tables(e);
        if (rc) efail(nullptr, rc, "mhd_ens_create: %s", e->err);
    }
    if (rc) {
        free_ens(e);
        delete e;
        return rc;
    }
    *out = e;
    return MHD_OK;
}

extern "C" int mhd_ens_set_metrics(mhd_ens_t *e, int b, const double *dxxdx, const double *d2xxdx2,
                                   const double *ddx, const double *dyydy, const double *d2yydy2, const double *ddy,
                                   const double *dzzdz, const double *d2zzdz2, const double *ddz,
                                   const double *rkapa, const double *dkapa) {
    int rc = check_box(e, b, "mhd_ens_set_metrics");
    if (rc) return rc;
    EnsBox &B = e->box[b];
    const double *src[11] = {dxxdx, d2xxdx2, ddx, dyydy, d2yydy2, ddy, dzzdz, d2zzdz2, ddz, rkapa, dkapa};
    size_t len[11];
    metric_lengths(B.p, len);
    const double *dst[11] = {e->hbox[b].m.dxxdx, e->hbox[b].m.d2xxdx2, e->hbox[b].m.ddx, e->hbox[b].m.dyydy,
                             e->hbox[b].m.d2yydy2, e->hbox[b].m.ddy, e->hbox[b].m.dzzdz, e->hbox[b].m.d2zzdz2,
                             e->hbox[b].m.ddz, e->hbox[b].m.rkapa, e->hbox[b].m.dkapa};
    EHIP(e, hipSetDevice(e->dev));
    for (int n = 0; n < 11; ++n) {
        if (!src[n]) return efail(e, MHD_ERR_ARG, "mhd_ens_set_metrics: array %d is null", n);
        EHIP(e, hipMemcpyAsync(const_cast<double *>(dst[n]), src[n], len[n] * sizeof(double), hipMemcpyHostToDevice,
                               e->stream));
    }
    // the grid-only CFL term, which the stage kernels take from here
    const double d2 = dd2_min(B.p, ddx, ddy, ddz);
    EHIP(e, hipMemcpyAsync(reinterpret_cast<char *>(e->sc + b) + offsetof(DevScalars, dd2min), &d2, sizeof(d2),
                           hipMemcpyHostToDevice, e->stream));
    const size_t nz = (size_t)B.p.npz + 2;
    B.dzzdz.assign(dzzdz, dzzdz + nz);
    B.d2zzdz2.assign(d2zzdz2, d2zzdz2 + nz);
    B.ddz.assign(ddz, ddz + nz);
    B.rkapa.assign(rkapa, rkapa + nz);
    B.dkapa.assign(dkapa, dkapa + nz);
    EHIP(e, hipStreamSynchronize(e->stream));
    B.have_metrics = true;
    return MHD_OK;
}

extern "C" int mhd_ens_set_background(mhd_ens_t *e, int b, const double *ro0, const double *tt0) {
    int rc = check_box(e, b, "mhd_ens_set_background");
    if (rc) return rc;
    EnsBox &B = e->box[b];
    if (!e->mixed) return efail(e, MHD_ERR_STATE, "mhd_ens_set_background: only for MHD_MIXED ensembles");
    if (!B.have_metrics) return efail(e, MHD_ERR_STATE, "mhd_ens_set_background: box %d: set the metrics first", b);
    if (!ro0 || !tt0) return efail(e, MHD_ERR_ARG, "mhd_ens_set_background: null array");
    std::vector<double> d;
    std::vector<float> f;
    std::vector<PlaneM> pl;
    const int bad = background_tables(B.p, B.dzzdz.data(), B.d2zzdz2.data(), B.ddz.data(), B.rkapa.data(),
                                      B.dkapa.data(), ro0, tt0, d, f, pl);
    if (bad >= 0) return efail(e, MHD_ERR_ARG, "mhd_ens_set_background: box %d: rho0[%d] <= 0", b, bad);
    EHIP(e, hipSetDevice(e->dev));
    EHIP(e, hipMemcpyAsync(const_cast<double *>(e->hbox[b].bg.ro), d.data(), d.size() * sizeof(double),
                           hipMemcpyHostToDevice, e->stream));
    EHIP(e, hipMemcpyAsync(const_cast<float *>(e->hbox[b].bg.ro32), f.data(), f.size() * sizeof(float),
                           hipMemcpyHostToDevice, e->stream));
    EHIP(e, hipMemcpyAsync(const_cast<PlaneM *>(e->hbox[b].bg.planes), pl.data(), pl.size() * sizeof(PlaneM),
                           hipMemcpyHostToDevice, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    B.have_bg = true;
    B.have_state = false;
    return MHD_OK;
}

extern "C" int mhd_ens_upload(mhd_ens_t *e, int b, const double *ru, const double *rv, const double *rw,
                              const double *ro, const double *tt, const double *bx, const double *by,
                              const double *bz) {
    int rc = check_box(e, b, "mhd_ens_upload");
    if (rc) return rc;
    EnsBox &B = e->box[b];
    const BoxDev &D = e->hbox[b];
    if (e->mixed && !B.have_bg) return efail(e, MHD_ERR_STATE, "mhd_ens_upload: box %d: set the background first", b);
    const double *src[8] = {ru, rv, rw, ro, tt, bx, by, bz};
    const int nf = e->lmag ? 8 : 5;
    const size_t pbytes = (size_t)(B.p.npx + 2) * (B.p.npy + 2) * (B.p.npz + 2) * sizeof(double);
    EHIP(e, hipSetDevice(e->dev));
    for (int f = 0; f < nf; ++f) {
        if (!src[f]) return efail(e, MHD_ERR_ARG, "mhd_ens_upload: field %d is null", f);
        EHIP(e, hipMemcpyAsync(e->stage, src[f], pbytes, hipMemcpyHostToDevice, e->stream));
        const size_t off = (size_t)f * B.p.fsize;
        if (!e->mixed) {
            EHIP(e, launch_pack(B.p, e->stage, D.q[0] + off, e->stream));
            EHIP(e, launch_pack(B.p, e->stage, D.q[1] + off, e->stream));
        } else {
            EHIP(e, launch_pack_mixed(B.p, e->stage, f, D.bg, D.h[1], D.rr, e->stream));
            EHIP(e, launch_pack_mixed(B.p, e->stage, f, D.bg, D.h[0], D.rr, e->stream));
        }
    }
    EHIP(e, hipStreamSynchronize(e->stream));
    B.have_state = true;
    B.cfl_valid = false;
    return MHD_OK;
}

extern "C" int mhd_ens_download(mhd_ens_t *e, int b, double *ru, double *rv, double *rw, double *ro, double *tt,
                                double *bx, double *by, double *bz) {
    int rc = check_box(e, b, "mhd_ens_download");
    if (rc) return rc;
    EnsBox &B = e->box[b];
    const BoxDev &D = e->hbox[b];
    if (!B.have_state) return efail(e, MHD_ERR_STATE, "mhd_ens_download: box %d: no state uploaded", b);
    double *dst[8] = {ru, rv, rw, ro, tt, bx, by, bz};
    const int nf = e->lmag ? 8 : 5;
    const size_t pbytes = (size_t)(B.p.npx + 2) * (B.p.npy + 2) * (B.p.npz + 2) * sizeof(double);
    EHIP(e, hipSetDevice(e->dev));
    for (int f = 0; f < nf; ++f) {
        if (!dst[f]) continue;
        const size_t off = (size_t)f * B.p.fsize;
        if (!e->mixed) EHIP(e, launch_unpack(B.p, D.q[e->cur] + off, e->stage, 1, e->stream));
        else EHIP(e, launch_unpack_mixed(B.p, D.h[e->cur], D.rr, f, D.bg, e->stage, e->stream));
        EHIP(e, hipMemcpyAsync(dst[f], e->stage, pbytes, hipMemcpyDeviceToHost, e->stream));
        EHIP(e, hipStreamSynchronize(e->stream));
    }
    return MHD_OK;
}

extern "C" int mhd_ens_set_clock(mhd_ens_t *e, int b, int nit, double timc, double timi) {
    int rc = check_box(e, b, "mhd_ens_set_clock");
    if (rc) return rc;
    DevScalars s;
    EHIP(e, hipSetDevice(e->dev));
    EHIP(e, hipMemcpyAsync(&s, e->sc + b, sizeof(s), hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    s.nit = nit;
    s.timc = timc;
    s.timi = timi;
    s.timt = timi + timc;
    EHIP(e, hipMemcpyAsync(e->sc + b, &s, sizeof(s), hipMemcpyHostToDevice, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    return MHD_OK;
}

extern "C" int mhd_ens_set_fixed_dt(mhd_ens_t *e, int b, double dt) {
    int rc = check_box(e, b, "mhd_ens_set_fixed_dt");
    if (rc) return rc;
    if (!(dt >= 0.0) || dt == INFINITY) return efail(e, MHD_ERR_ARG, "mhd_ens_set_fixed_dt: dt=%g", dt);
    DevScalars s;
    EHIP(e, hipSetDevice(e->dev));
    EHIP(e, hipMemcpyAsync(&s, e->sc + b, sizeof(s), hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    s.fixed_dt = dt;
    if (dt > 0.0) s.dt = dt;
    else e->box[b].cfl_valid = false;
    EHIP(e, hipMemcpyAsync(e->sc + b, &s, sizeof(s), hipMemcpyHostToDevice, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    return MHD_OK;
}

extern "C" int mhd_ens_get_clock(mhd_ens_t *e, int b, int *nit, double *timc, double *timt) {
    int rc = check_box(e, b, "mhd_ens_get_clock");
    if (rc) return rc;
    DevScalars s;
    EHIP(e, hipSetDevice(e->dev));
    EHIP(e, hipMemcpyAsync(&s, e->sc + b, sizeof(s), hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    if (nit) *nit = s.nit;
    if (timc) *timc = s.timc;
    if (timt) *timt = s.timt;
    return MHD_OK;
}

// Launches of one ensemble step reading register a (direct or captured).
static hipError_t enqueue_step(mhd_ens_t *e, int a) {
    hipError_t r = hipSuccess;
    const int grid = e->grid > 0 && e->grid < e->nitems ? e->grid : e->nitems;
    for (int st = 1; st <= 3 && r == hipSuccess; ++st) {
        if (!e->mixed) {
            r = launch_stage_batch(st, e->tile_y, e->lmag, e->dbox, e->ditems, e->nitems, grid, a, e->stream);
            if (r == hipSuccess) r = launch_bcon_batch(e->lmag, e->dbox, e->dcols, e->ncols, a, st == 3, e->stream);
        } else {
            r = launch_stage_batch_mixed(st, e->tile_x, e->tile_y, e->lmag, e->dbox, e->ditems, e->nitems, grid, a,
                                         e->stream);
            if (r == hipSuccess)
                r = launch_bcon_batch_mixed(e->lmag, e->dbox, e->dcols, e->ncols, a, st == 3, e->stream);
        }
        a = 1 - a;
    }
    if (r == hipSuccess) r = launch_cfl_final_batch(e->dbox, e->nbox, true, e->stream);
    return r;
}

// Stand-alone CFL reduction of the boxes uploaded since their last step.
static int prime_cfl(mhd_ens_t *e) {
    for (int b = 0; b < e->nbox; ++b) {
        EnsBox &B = e->box[b];
        if (!B.have_metrics || !B.have_state)
            return efail(e, MHD_ERR_STATE, "box %d: metrics and state must be set before stepping", b);
        if (B.cfl_valid) continue;
        const BoxDev &D = e->hbox[b];
        if (!e->mixed) EHIP(e, launch_cfl(B.p, D.m, D.q[e->cur], D.sc, e->stream));
        else EHIP(e, launch_cfl_mixed(B.p, D.m, D.bg, D.h[e->cur], D.rr, D.sc, e->stream));
        EHIP(e, launch_cfl_final(B.p, D.sc, false, e->stream));
        B.cfl_valid = true;
    }
    return MHD_OK;
}

static int get_graph(mhd_ens_t *e, int pair, hipGraphExec_t *out) {
    hipGraphExec_t &g = e->gexec[pair][e->cur];
    if (!g) {
        hipGraph_t graph = nullptr;
        EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        hipError_t r = enqueue_step(e, e->cur);
        if (r == hipSuccess && pair) r = enqueue_step(e, 1 - e->cur);
        hipError_t r2 = hipStreamEndCapture(e->stream, &graph);
        if (r != hipSuccess || r2 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return efail(e, MHD_ERR_HIP, "graph capture: %s",
                         hipGetErrorString(r != hipSuccess ? r : r2));
        }
        r = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (r != hipSuccess) {
            g = nullptr;
            return efail(e, MHD_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(r));
        }
    }
    *out = g;
    return MHD_OK;
}

// n steps, queued on the stream (no synchronisation).
static int run_steps(mhd_ens_t *e, int n, int use_graph) {
    int rc = prime_cfl(e);
    if (rc) return rc;
    int done = 0;
    if (use_graph) {
        // the first launch of each kernel variant sets its LDS attribute, which
        // must not happen inside a capture: take one direct step first
        if (!e->gexec[0][0] && !e->gexec[0][1] && !e->gexec[1][0] && !e->gexec[1][1] && n > 0) {
            EHIP(e, enqueue_step(e, e->cur));
            e->cur = 1 - e->cur;
            done = 1;
        }
        hipGraphExec_t g;
        if (n - done >= 2) {
            rc = get_graph(e, 1, &g);
            if (rc) return rc;
            for (; done + 2 <= n; done += 2) EHIP(e, hipGraphLaunch(g, e->stream));
        }
        if (done < n) {
            rc = get_graph(e, 0, &g);
            if (rc) return rc;
            EHIP(e, hipGraphLaunch(g, e->stream));
            e->cur = 1 - e->cur;
            ++done;
        }
    } else {
        for (; done < n; ++done) {
            EHIP(e, enqueue_step(e, e->cur));
            e->cur = 1 - e->cur;
        }
    }
    return MHD_OK;
}

extern "C" int mhd_ens_steps(mhd_ens_t *e, int n, int use_graph, double *dt, double *umach) {
    if (!e) return efail(nullptr, MHD_ERR_ARG, "mhd_ens_steps: null handle");
    if (n < 0) return efail(e, MHD_ERR_ARG, "mhd_ens_steps: n=%d", n);
    EHIP(e, hipSetDevice(e->dev));
    int rc = run_steps(e, n, use_graph);
    if (rc) return rc;
    std::vector<DevScalars> s(e->nbox);
    EHIP(e, hipMemcpyAsync(s.data(), e->sc, e->nbox * sizeof(DevScalars), hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    int bad = -1;
    for (int b = 0; b < e->nbox; ++b) {
        if (dt) dt[b] = s[b].dt_used;
        if (umach) umach[b] = s[b].umach_used;
        if (s[b].nonfinite && bad < 0) bad = b;
    }
    if (bad >= 0)
        return efail(e, MHD_ERR_NONFINITE, "box %d: non-finite CFL limit or Mach number (DT=%g)", bad,
                     s[bad].dt_used);
    return MHD_OK;
}

extern "C" int mhd_ens_set_tile(mhd_ens_t *e, int tile_x, int tile_y, int kdepth, int grid) {
    if (!e) return efail(nullptr, MHD_ERR_ARG, "mhd_ens_set_tile: null handle");
    const bool shape_ok = (tile_x == 64 && (tile_y == 4 || tile_y == 8)) || (tile_x == 32 && tile_y == 8 && e->mixed);
    if (!shape_ok || kdepth < 0 || grid < 0)
        return efail(e, MHD_ERR_ARG,
                     "mhd_ens_set_tile: tile %dx%d (64x4, 64x8, or 32x8 in the mixed form), kdepth %d, grid %d",
                     tile_x, tile_y, kdepth, grid);
    EHIP(e, hipSetDevice(e->dev));
    EHIP(e, hipStreamSynchronize(e->stream));
    e->tile_x = tile_x;
    e->tile_y = tile_y;
    e->kdepth = kdepth;
    e->grid = grid;
    return build_tables(e);
}

extern "C" int mhd_ens_get_tile(const mhd_ens_t *e, int *tile_x, int *tile_y, int *kdepth, int *grid) {
    if (!e) return MHD_ERR_ARG;
    if (tile_x) *tile_x = e->tile_x;
    if (tile_y) *tile_y = e->tile_y;
    if (kdepth) *kdepth = e->kdepth;
    if (grid) *grid = e->grid;
    return MHD_OK;
}

extern "C" int mhd_ens_set_tiling(mhd_ens_t *e, int tile_y, int kdepth, int grid) {
    if (!e) return efail(nullptr, MHD_ERR_ARG, "mhd_ens_set_tiling: null handle");
    if ((tile_y != 4 && tile_y != 8) || kdepth < 0 || grid < 0)
        return efail(e, MHD_ERR_ARG, "mhd_ens_set_tiling: tile_y %d (4 or 8), kdepth %d, grid %d", tile_y, kdepth,
                     grid);
    return mhd_ens_set_tile(e, 64, tile_y, kdepth, grid);
}

extern "C" int mhd_ens_get_tiling(const mhd_ens_t *e, int *tile_y, int *kdepth, int *grid) {
    return mhd_ens_get_tile(e, nullptr, tile_y, kdepth, grid);
}

// Mean ms per step over n steps queued with or without graphs.
static int time_run(mhd_ens_t *e, int n, int use_graph, double *ms) {
    hipEvent_t a, b;
    EHIP(e, hipEventCreate(&a));
    hipError_t r = hipEventCreate(&b);
    if (r != hipSuccess) {
        (void)hipEventDestroy(a);
        return efail(e, MHD_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(r));
    }
    int rc = MHD_OK;
    r = hipEventRecord(a, e->stream);
    if (r == hipSuccess) rc = run_steps(e, n, use_graph);
    if (r == hipSuccess && !rc) r = hipEventRecord(b, e->stream);
    if (r == hipSuccess && !rc) r = hipEventSynchronize(b);
    float t = 0.f;
    if (r == hipSuccess && !rc) r = hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (r != hipSuccess) return efail(e, MHD_ERR_HIP, "timing: %s", hipGetErrorString(r));
    *ms = t / n;
    return MHD_OK;
}

extern "C" int mhd_ens_time_steps(mhd_ens_t *e, int n, double *ms_step) {
    if (!e) return efail(nullptr, MHD_ERR_ARG, "mhd_ens_time_steps: null handle");
    if (n < 1) return efail(e, MHD_ERR_ARG, "mhd_ens_time_steps: n=%d", n);
    EHIP(e, hipSetDevice(e->dev));
    double ms = 0;
    int rc = time_run(e, n, 1, &ms);
    if (rc) return rc;
    if (ms_step) *ms_step = ms;
    return MHD_OK;
}

extern "C" int mhd_ens_autotune(mhd_ens_t *e, int nsteps, char *report, int report_len) {
    if (!e) return efail(nullptr, MHD_ERR_ARG, "mhd_ens_autotune: null handle");
    if (nsteps < 1) return efail(e, MHD_ERR_ARG, "mhd_ens_autotune: nsteps=%d", nsteps);
    EHIP(e, hipSetDevice(e->dev));
    int rc = prime_cfl(e);
    if (rc) return rc;
    // save what the next step reads: the current register of every box (and
    // the residual register R for the mixed form) and the device scalars; Z
    // or D and the other register are rewritten by the first stage
    const size_t esz = e->mixed ? sizeof(float) : sizeof(double);
    const int nf = e->lmag ? 8 : 5;
    size_t tot = 0;
    for (auto &B : e->box) tot += B.reg_elems * esz + (e->mixed ? resid_bytes(nf, B.p.fsize) : 0);
    char *save = nullptr;
    DevScalars *ssave = nullptr;
    EHIP(e, hipMalloc(&save, tot));
    hipError_t r = hipMalloc(&ssave, e->nbox * sizeof(DevScalars));
    if (r != hipSuccess) {
        (void)hipFree(save);
        return efail(e, MHD_ERR_HIP, "mhd_ens_autotune: %s", hipGetErrorString(r));
    }
    const int cur0 = e->cur;
    auto copy_state = [&](bool to_save) -> hipError_t {
        size_t off = 0;
        hipError_t x = hipSuccess;
        for (int b = 0; b < e->nbox && x == hipSuccess; ++b) {
            const BoxDev &D = e->hbox[b];
            const size_t nbs[2] = {e->box[b].reg_elems * esz, e->mixed ? resid_bytes(nf, e->box[b].p.fsize) : 0};
            void *regs[2] = {e->mixed ? (void *)D.h[cur0] : (void *)D.q[cur0], (void *)D.rr};
            for (int i = 0; i < (e->mixed ? 2 : 1) && x == hipSuccess; ++i) {
                x = to_save ? hipMemcpyAsync(save + off, regs[i], nbs[i], hipMemcpyDeviceToDevice, e->stream)
                            : hipMemcpyAsync(regs[i], save + off, nbs[i], hipMemcpyDeviceToDevice, e->stream);
                off += nbs[i];
            }
        }
        if (x == hipSuccess)
            x = to_save ? hipMemcpyAsync(ssave, e->sc, e->nbox * sizeof(DevScalars), hipMemcpyDeviceToDevice, e->stream)
                        : hipMemcpyAsync(e->sc, ssave, e->nbox * sizeof(DevScalars), hipMemcpyDeviceToDevice,
                                         e->stream);
        return x;
    };
    r = copy_state(true);
    const int t0x = e->tile_x, t0y = e->tile_y, t0k = e->kdepth, t0g = e->grid;
    const int txs[3] = {64, 64, 32}, tys[3] = {4, 8, 8};  // 32x8 in the mixed form only
    const int nshape = e->mixed ? 3 : 2;
    constexpr int NKD = 7;
    const int kds[NKD] = {8, 12, 16, 20, 24, 32, 0};
    const int per_cu[4] = {0, 1, 2, 4};  // 0: one workgroup per item
    double best = 1e300;
    int btx = t0x, bty = t0y, bkd = t0k, bgr = t0g;
    std::string rep = "[";
    for (int a = 0; a < nshape && r == hipSuccess && !rc; ++a)
        for (int k = 0; k < NKD && r == hipSuccess && !rc; ++k)
            for (int g = 0; g < 4 && r == hipSuccess && !rc; ++g) {
                const int grid = per_cu[g] * e->cus;
                rc = mhd_ens_set_tile(e, txs[a], tys[a], kds[k], grid);
                if (rc) break;
                double ms = 0;
                rc = time_run(e, 1, 0, &ms);  // warm-up (kernel attributes, caches)
                if (!rc) rc = time_run(e, nsteps, 0, &ms);
                if (rc) break;
                r = copy_state(false);
                e->cur = cur0;
                char line[160];
                snprintf(line, sizeof line,
                         "%s{\"tile_x\": %d, \"tile_y\": %d, \"kdepth\": %d, \"grid\": %d, \"ms_step\": %.5f}",
                         rep.size() > 1 ? ", " : "", txs[a], tys[a], kds[k], grid, ms);
                rep += line;
                if (ms < best) {
                    best = ms;
                    btx = txs[a];
                    bty = tys[a];
                    bkd = kds[k];
                    bgr = grid;
                }
            }
    rep += "]";
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    (void)hipFree(save);
    (void)hipFree(ssave);
    if (rc) return rc;
    if (r != hipSuccess) return efail(e, MHD_ERR_HIP, "mhd_ens_autotune: %s", hipGetErrorString(r));
    rc = mhd_ens_set_tile(e, btx, bty, bkd, bgr);
    if (rc) return rc;
    if (report && report_len > 0) {
        strncpy(report, rep.c_str(), (size_t)report_len - 1);
        report[report_len - 1] = 0;
    }
    return MHD_OK;
}

extern "C" long long mhd_ens_zones(const mhd_ens_t *e) {
    if (!e) return 0;
    long long z = 0;
    for (auto &B : e->box) z += (long long)B.p.npx * B.p.npy * B.p.npz;
    return z;
}

extern "C" const char *mhd_ens_last_error(const mhd_ens_t *e) { return e ? e->err : g_ens_err; }

extern "C" void mhd_ens_destroy(mhd_ens_t *e) {
    if (!e) return;
    free_ens(e);
    delete e;
}
