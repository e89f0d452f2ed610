// Device-side data structures shared by the kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

struct mhd_net;

namespace rmhd {

// A kernel's dynamic-LDS limit, set once per device: function attributes
// belong to the device current at the call, so a process that drives
// several GPUs sets them on each.  done: one bit per device index (< 64).
inline hipError_t lds_limit_once(std::atomic<unsigned long long> &done, const void *fn, int bytes) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const unsigned long long bit = (dev >= 0 && dev < 64) ? 1ull << dev : 0ull;
    if (bit && (done.load(std::memory_order_acquire) & bit)) return hipSuccess;
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess && bit) done.fetch_or(bit, std::memory_order_acq_rel);
    return e;
}

// Field order in every device register (q, Z, F).
enum Field { F_RU = 0, F_RV, F_RW, F_RO, F_TT, F_BX, F_BY, F_BZ, NFIELD };

// Elements between the end of one field's planes and the start of the
// next: 37 lines of 128 B in FP32 (74 in FP64), so that the fields of one
// point do not all fall into the same L2 set, as they do when the field
// stride is a multiple of the plane size of a power-of-two box (the stage
// kernels then re-read halo rows that such conflicts had evicted).
constexpr long long FIELD_PAD = 1184;

// Scalars of the step, all as the driver computed them, plus products the
// reference forms once per FLUXES call (e.g. TMP=OCV*ORE), computed on the
// host with the same association.
struct Params {
    int npx, npy, npz;          // interior sizes
    int lmag, lrot, izc, itc, ibc;
    int tile_y;                 // rows per stage-kernel tile (4 or 8)
    long long plane;            // npx*npy
    long long fsize;            // field stride: plane*(npz+2) values incl. z ghost planes, + FIELD_PAD
    double hx, h2x, hy, h2y, hz, h2z, c13, c23, c43;
    double gam1, gam2, gam3, zeta1, zeta2;
    double cv, ocv, ore, re, repr, theta, grav, sf, gamma;
    double omx, omz, orm, rm, obeta, tb, dzf;
    double ogamma;              // 1/GAMMA (sub:744)
    double t_cond;              // OCV/REPR (sub:1584)
    double t_visc;              // OCV*ORE (sub:1904)
    double t_visc43;            // C43*OCV*ORE (sub:1962)
    double t_cross;             // C13*ORE (sub:1930)
    double t_ohm;               // OBETA*ORM*OCV (sub:2287)
    // Optional branches of FLUXES/BCON (FP64 form; all zero: the main path).
    int opt;                    // any branch below that the stage kernel must see
    int id, lrem, lshr, mirror; // mirror: ID or LREM (ghost RO planes mirrored, sub:1301-1310)
    int npey, npez;             // decomposition whose summation order the means follow
    double dh, dv, tu, dztu, dztb, rlax, tstart, toff, tp, tc, hh;
    double ramp;                // 1 + TANH(LOG(3.0E00)/QFH*(TIMT-TC)) of this step (ITC=2, TC != 0)
    double tpr;                 // plume TPR of this step (TP with ITC=0/1)
    // per-plane (npz+2) and per-axis device tables of the branches
    const double *rom, *wwz, *ttm, *hrad, *radd;
    const double *gx, *gy, *gz; // EXP(CLN*(EXX(I)-XP)**2) etc. per padded index
    double *ftacc;              // LSHR: FT carried between evaluations (one field)
};

// Device metric table: padded arrays, same indexing as the reference
// (element I of the reference is at index I-1).
struct Metrics {
    const double *dxxdx, *d2xxdx2, *ddx;
    const double *dyydy, *d2yydy2, *ddy;
    const double *dzzdz, *d2zzdz2, *ddz;
    const double *rkapa, *dkapa;
    const double *rkcfl;        // conductivity of the CFL limit: RKAPA, or 1+RKAPA with LREM (sub:844-852)
};

// Per-step device scalars.  The reductions keep order-preserving integer
// keys of the doubles so that min/max can use 64-bit integer atomics.
struct DevScalars {
    double dt, umach;                 // for the next step (from the last reduction)
    double dt_used, umach_used;       // of the step that last completed
    double cflmin[4];
    double timc, timi, timt;
    unsigned long long key_mach;      // max of OGAMMA*UU/TT
    unsigned long long key_min[4];    // mins of the CFL limits
    double fixed_dt;                  // > 0: step with this dt instead of the CFL dt
    double dd2min;                    // min over all points of min(DDX,DDY,DDZ)^2 (metric only)
    int nit;
    int nonfinite;
};

// Hydrostatic background of the mixed form, per plane k = 0..npz+1:
// FP64 rho0, T0 and the background's own discrete residuals of the momentum
// z equation and the temperature equation; FP32 copies and per-plane
// coefficients for the perturbation stencil.
// Everything the mixed stage kernel reads per plane (metrics and
// background), as one 160-byte record: the z march holds one table pointer
// instead of one per array, and reads the fields it needs through scalar
// loads where it uses them.  The FP32 coefficients of the right-hand side
// (the same in every record) are read there too: loads at plane-dependent
// addresses stay in the plane instead of holding scalar registers for the
// whole march.
struct PlaneM {
    double dzzdz, d2zzdz2, rkapa, dkapa, ddz, ro, tt, fwbg, ftbg, iddz2, irk;
    float ro32, tt32, dt0dz32, rinv0_32, s0_32;
    float c43, ocv, ore, orm, obeta, grav, t_cond, t_visc, t_visc43, t_cross, t_ohm, omx, omz;
};
static_assert(sizeof(PlaneM) == 160, "PlaneM is one 160-byte record");

struct Background {
    const double *ro, *tt, *fwbg, *ftbg;
    const double *iddz2, *irk;  // 1/ddz^2 and 1/rkapa per plane (grid only; the CFL of stage 3)
    const PlaneM *planes;       // the same per plane with the metrics, for the stage kernel
    const float *ro32, *tt32, *dt0dz32, *rinv0_32, *s0_32;
    // optional subgrid forcing of the closure-corrected step: FP32 fields
    // [RU,RV,RW,TT,BX,BY,BZ][plane][j][i] added to the right-hand side of
    // every stage while *force_on != 0 (null: none)
    const float *force;
    const int *force_on;
};

// One box of an ensemble as the batched kernels see it (device memory).
// FP64 boxes use q/z, mixed boxes h/dr/rr and bg.
struct BoxDev {
    Params p;
    Metrics m;
    Background bg;
    double *q[2];
    double *z;
    float *h[2];
    float *dr;
    unsigned *rr;
    DevScalars *sc;
};

// Bytes of the mixed form's residual register R for nf fields.
inline size_t resid_bytes(int nf, long long fsize) { return (size_t)((nf + 1) / 2) * (size_t)fsize * 4; }

// Work item of a batched stage launch: one tile column chunk of one box.
struct StageItem {
    int box, i0, j0, kbeg, kend;
};

// Work item of a batched boundary-plane launch: 256 columns of one box.
struct ColumnItem {
    int box, col0;
};

__host__ __device__ inline unsigned long long dkey(double v) {
    unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
// Updates of a reduction key.  Between its reset (by an earlier launch or
// the host) and its read (by a later launch) a key only moves one way, so
// a workgroup whose value cannot move it, by a read of the key at device
// scope (coherent across the XCDs' L2 caches), skips the atomic: a stale
// read can only cost an atomic that was not needed, never skip one that
// was.  Without the check every workgroup of a launch queues an atomic on
// the same cache line; for the 1024 short workgroups of a boundary-plane
// launch that queue was three quarters of the launch.
__device__ inline void key_raise(unsigned long long *k, unsigned long long v) {
    if (v > __hip_atomic_load(k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(k, v);
}
__device__ inline void key_lower(unsigned long long *k, unsigned long long v) {
    if (v < __hip_atomic_load(k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(k, v);
}
__host__ __device__ inline double dunkey(unsigned long long k) {
    unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __builtin_bit_cast(double, b);
}

// Saved state of the flux relaxation (FLUXES' DATA/SAVE, sub:1241-1245).
struct RelaxState {
    double tfac, ttm_max, ttm_min;
    int icall;
};

// Device buffers of the optional branches: the per-plane tables the stage
// kernel reads (npz+2 each), slab partial sums (2*(npz+2)*npey), relaxation
// work space (3*(npz+2)+npez), the relaxation state and the coordinates ZEE.
struct OptBuffers {
    double *rom, *wwz, *ttm, *hrad, *radd;
    double *part, *work;
    RelaxState *rs;
    const double *zee;
};
inline size_t opt_part_len(const Params &p) { return 2 * (size_t)(p.npz + 2) * p.npey; }
inline size_t opt_work_len(const Params &p) { return 3 * (size_t)(p.npz + 2) + p.npez; }

// Before every FLUXES evaluation of the optional branches (mhd_opt.hip), from
// the stage input q: the plane means ROM and TTM (HORIZONTAL_MEAN in the
// summation order of npey slabs), WWZ of ID=1, HRAD of LREM, and the
// relaxation increment radd = RLAX*TFAC*CORRECT with its saved state.
hipError_t launch_opt_means(const Params &p, const Metrics &m, const double *q, const OptBuffers &o,
                            DevScalars *sc, hipStream_t s);

// Launchers (mhd_kernels.hip).
hipError_t launch_stage(int mode, const Params &p, const Metrics &m, const double *qin, double *qout,
                        double *z, double *fout, DevScalars *sc, hipStream_t s);
hipError_t launch_bcon(const Params &p, const Metrics &m, const double *qin, double *qout, DevScalars *sc,
                       bool cfl, hipStream_t s);
hipError_t launch_cfl(const Params &p, const Metrics &m, const double *q, DevScalars *sc, hipStream_t s);
hipError_t launch_cfl_final(const Params &p, DevScalars *sc, bool clock, hipStream_t s);
hipError_t launch_pack(const Params &p, const double *padded, double *field, hipStream_t s);
hipError_t launch_unpack(const Params &p, const double *field, double *padded, int mode, hipStream_t s);

// Ensemble launchers (all boxes share lmag and the tile height).  Register
// a of every box is the stage input.  grid is the persistent grid size.
hipError_t launch_stage_batch(int mode, int tile_y, int lmag, const BoxDev *boxes, const StageItem *items,
                              int nitems, int grid, int a, hipStream_t s);
hipError_t launch_bcon_batch(int lmag, const BoxDev *boxes, const ColumnItem *items, int nitems, int a, bool cfl,
                             hipStream_t s);
hipError_t launch_cfl_final_batch(const BoxDev *boxes, int nbox, bool clock, hipStream_t s);
hipError_t launch_box_filter(const Params &p, const double *q, int nf, int factor, int zmode, int nxc, int nyc, int nzc,
                             double *out, hipStream_t s);

// Closure-corrected mixed step (mhd_mixed.hip): network inputs from the
// state (rho, u, v, w, s = ln T/(gamma-1) - ln rho, BX, BY, BZ; nrm =
// per-plane means [8][npz], 1/std[8]) as BF16 [plane-1][j][i][8], and the
// forcing from the network output (nrm_out = per-plane mean subgrid terms
// [7][npz], std[7]); force_on is cleared if any output is not finite.
hipError_t launch_closure_inputs(const Params &p, const Background &bg, const float *h, const unsigned *rr,
                                 const float *nrm, __bf16 *x, hipStream_t s);
hipError_t launch_closure_forcing(const Params &p, const float *y, const float *nrm_out, float *force, int *force_on,
                                  hipStream_t s);

// Horizontal means and mean convective flux per interior plane (mhd_diag.hip):
// out[3][npz] = TTM, ROM, FCONM.
hipError_t launch_flux_profiles(const Params &p, const double *q, double *out, hipStream_t s);
hipError_t launch_flux_profiles_mixed(const Params &p, const float *h, const unsigned *rr, const Background &bg,
                                      double *out, hipStream_t s);

// Mixed-precision form (mhd_mixed.hip).  Registers: h (FP32 stored values,
// ping-pong), dr (FP32: residual + zeta dt F between stages 1-2 and 2-3) and
// rr (the rounding residual of the stored values between steps as BF16, two
// fields per 32-bit word: (nf+1)/2 words per point, written by stage 3 and
// the upload, read by stage 1).
hipError_t launch_stage_mixed(int mode, const Params &p, const Metrics &m, const Background &bg, const float *hin,
                              float *hout, float *dr, unsigned *rr, DevScalars *sc, hipStream_t s);
hipError_t launch_bcon_mixed(const Params &p, const Metrics &m, const Background &bg, float *h, float *dr,
                             unsigned *rr, DevScalars *sc, bool last, hipStream_t s);
hipError_t launch_cfl_mixed(const Params &p, const Metrics &m, const Background &bg, const float *h,
                            const unsigned *rr, DevScalars *sc, hipStream_t s);
// Conversion of field `field` between a padded FP64 array and the stored
// values / resting residuals of the registers hreg, rreg (base pointers, all
// fields).  Pack rho (field 3) before T (field 4).
hipError_t launch_pack_mixed(const Params &p, const double *padded, int field, const Background &bg, float *hreg,
                             unsigned *rreg, hipStream_t s);
hipError_t launch_unpack_mixed(const Params &p, const float *hreg, const unsigned *rreg, int field,
                               const Background &bg, double *padded, hipStream_t s);
// Re-expresses rho and T of the interior planes about the background nb
// instead of ob, in place (mhd_rebase).
hipError_t launch_rebase_mixed(const Params &p, const Background &ob, const Background &nb, float *hreg,
                               unsigned *rreg, hipStream_t s);

hipError_t launch_stage_batch_mixed(int mode, int tile_x, int tile_y, int lmag, const BoxDev *boxes,
                                    const StageItem *items, int nitems, int grid, int a, hipStream_t s);
hipError_t launch_bcon_batch_mixed(int lmag, const BoxDev *boxes, const ColumnItem *items, int nitems, int a,
                                   bool last, hipStream_t s);

// Network entry points used by the closure-corrected step (mhd_net.hip).
int net_shape(const mhd_net *n, int *cin, int *cout);
int net_prepare(mhd_net *n, long long nvox, int *dev);
hipError_t net_forward_device(mhd_net *n, hipStream_t s, int nx, int ny, int nz, const __bf16 *x, const float **y);

}  // namespace rmhd
