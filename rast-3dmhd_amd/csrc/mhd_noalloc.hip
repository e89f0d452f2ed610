// mhd_step_noalloc (include/rast3dmhd.h): one plain FP64 step on arrays the
// caller keeps on the device, with no device allocation in the call.
//
// The step's working registers (q[0], q[1], Z) and its reduction scalars are
// the scratch of the handle-based step (mhd_capi.hip allocates them in
// mhd_create).  Here they live in one file-scope device array, `scratch`,
// of NOALLOC_BYTES per GPU:
//
//   [0, 256)             DevScalars (reduction keys, DT, clock)
//   [256, 256 + 3*R)     q[0], q[1], Z; R = nf * fsize doubles each (fsize
//                        the field stride of Params, padding included)
//
// On the host the name `scratch` stands for no device address: its address
// on each device is looked up once per device with hipGetSymbolAddress and
// only that pointer is passed to kernels.  Every call re-initialises the
// scalars with a one-thread kernel and overwrites the registers completely
// (pack of every field, z ghost planes included), so no content of an
// earlier call is read.  Calls that overlap on one GPU share the array and
// race.
//
// The launches are those of mhd_upload + mhd_step + mhd_download on a
// handle (the same kernels with the same arguments), so the result is the
// handle's bit for bit.

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

#include "../../include/rast3dmhd.h"
#include "mhd_host.h"

using namespace rmhd;

namespace {

constexpr size_t NOALLOC_BYTES = 64ull << 20;  // per GPU, held until the process ends
constexpr size_t SCALARS_BYTES = 256;
static_assert(sizeof(DevScalars) <= SCALARS_BYTES, "DevScalars fits its 256-byte slot");

__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEV = 64;
std::mutex g_mu;
void *g_scratch[MAX_DEV];  // device address of `scratch` per device (null: not looked up yet)

__global__ void k_noalloc_init(DevScalars *sc, DevScalars s0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *sc = s0;
}

// Device address of this device's scratch array (dev is current).
hipError_t scratch_ptr(int dev, unsigned char **out) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = p;
    }
    *out = static_cast<unsigned char *>(g_scratch[dev]);
    return hipSuccess;
}

}  // namespace

#define NACHK(x)                                                                                     \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            return global_fail(MHD_ERR_HIP, "mhd_step_noalloc: %s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

extern "C" int mhd_step_noalloc(const mhd_desc *d, const mhd_dev_metrics *met, double *const *state, int *nit,
                                double *timc, double *timt, double timi, double *dt, double *umach, void *stream) {
    if (!d || !met || !state || !nit || !timc || !timt)
        return global_fail(MHD_ERR_ARG, "mhd_step_noalloc: null argument");
    {
        char msg[400];
        const int rc = check_desc(d, msg, sizeof msg);
        if (rc) return global_fail(rc, "mhd_step_noalloc: %s", msg);
    }
    if (d->precision != MHD_FP64) return global_fail(MHD_ERR_SCOPE, "mhd_step_noalloc: MHD_FP64 only");
    if (desc_branches(d))
        return global_fail(MHD_ERR_SCOPE, "mhd_step_noalloc: the main path only (no optional FLUXES/BCON branch)");
    const Params p = params_from_desc(d);
    const int nf = p.lmag ? 8 : 5;
    const size_t reg = (size_t)nf * (size_t)p.fsize;  // doubles per register
    const size_t need = SCALARS_BYTES + 3 * reg * sizeof(double);
    if (need > NOALLOC_BYTES)
        return global_fail(MHD_ERR_ARG,
                           "mhd_step_noalloc: box %dx%dx%d needs %zu bytes of scratch, the maximum is %zu "
                           "(use a handle: mhd_create)",
                           d->npx, d->npy, d->npz, need, NOALLOC_BYTES);
    const double *marr[11] = {met->dxxdx, met->d2xxdx2, met->ddx, met->dyydy, met->d2yydy2, met->ddy,
                              met->dzzdz, met->d2zzdz2, met->ddz, met->rkapa, met->dkapa};
    for (int n = 0; n < 11; ++n)
        if (!marr[n]) return global_fail(MHD_ERR_ARG, "mhd_step_noalloc: metric array %d is null", n);
    for (int f = 0; f < nf; ++f)
        if (!state[f]) return global_fail(MHD_ERR_ARG, "mhd_step_noalloc: field %d is null", f);

    int dev = 0;
    NACHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEV) return global_fail(MHD_ERR_HIP, "mhd_step_noalloc: device %d", dev);
    unsigned char *base = nullptr;
    NACHK(scratch_ptr(dev, &base));
    DevScalars *sc = reinterpret_cast<DevScalars *>(base);
    double *q[2] = {reinterpret_cast<double *>(base + SCALARS_BYTES),
                    reinterpret_cast<double *>(base + SCALARS_BYTES) + reg};
    double *z = q[1] + reg;
    const hipStream_t s = static_cast<hipStream_t>(stream);

    Metrics m;
    m.dxxdx = met->dxxdx; m.d2xxdx2 = met->d2xxdx2; m.ddx = met->ddx;
    m.dyydy = met->dyydy; m.d2yydy2 = met->d2yydy2; m.ddy = met->ddy;
    m.dzzdz = met->dzzdz; m.d2zzdz2 = met->d2zzdz2; m.ddz = met->ddz;
    m.rkapa = met->rkapa; m.dkapa = met->dkapa;
    m.rkcfl = met->rkapa;

    // scalars: keys armed, the caller's clock
    DevScalars s0 = initial_scalars();
    s0.nit = *nit;
    s0.timc = *timc;
    s0.timi = timi;
    s0.timt = timi + *timc;
    hipLaunchKernelGGL(k_noalloc_init, dim3(1), dim3(64), 0, s, sc, s0);
    NACHK(hipGetLastError());
    // upload: both registers carry the z ghost planes (as mhd_upload)
    for (int f = 0; f < nf; ++f) {
        NACHK(launch_pack(p, state[f], q[0] + (size_t)f * p.fsize, s));
        NACHK(launch_pack(p, state[f], q[1] + (size_t)f * p.fsize, s));
    }
    // step (as mhd_step on a freshly uploaded handle)
    NACHK(launch_cfl(p, m, q[0], sc, s));
    NACHK(launch_cfl_final(p, sc, false, s));
    int a = 0, b = 1;
    for (int st = 1; st <= 3; ++st) {
        NACHK(launch_stage(st, p, m, q[a], q[b], z, nullptr, sc, s));
        NACHK(launch_bcon(p, m, q[a], q[b], sc, st == 3, s));
        const int t = a;
        a = b;
        b = t;
    }
    NACHK(launch_cfl_final(p, sc, true, s));
    // download into the caller's arrays (as mhd_download), register a holds the result
    for (int f = 0; f < nf; ++f) NACHK(launch_unpack(p, q[a] + (size_t)f * p.fsize, state[f], 1, s));
    DevScalars out;
    NACHK(hipMemcpyAsync(&out, sc, sizeof(out), hipMemcpyDeviceToHost, s));
    NACHK(hipStreamSynchronize(s));
    *nit = out.nit;
    *timc = out.timc;
    *timt = out.timt;
    if (dt) *dt = out.dt_used;
    if (umach) *umach = out.umach_used;
    if (out.nonfinite)
        return global_fail(MHD_ERR_NONFINITE, "mhd_step_noalloc: non-finite CFL limit or Mach number (DT=%g, UMACH=%g)",
                           out.dt_used, out.umach_used);
    return MHD_OK;
}
