/*
 * rast3dmhd.h -- C-ABI of the MI355X explicit compressible-MHD time step.
 *
 * This is the drop-in boundary for the reference's SUBROUTINE STEP
 * (3dmhdsub.f:687-1176), which takes no arguments and exchanges everything
 * through COMMON blocks (3dmhd.f:51-71).  A Fortran STEP shim with the same
 * name and COMMONs (rast-3dmhd_amd/fortran/step_shim.f) binds these entry
 * points with ISO_C_BINDING; the driver is unchanged.
 *
 * Conventions
 *  - Every array argument is a host pointer to FP64 data in the reference's
 *    padded Fortran order: 3-D fields are (NX,NY,NZ) = (NPX+2,NPY+2,NPZ+2),
 *    x fastest, one ghost layer per side; metric arrays have length NX, NY
 *    or NZ.  The library never keeps a host pointer past a call.
 *  - Every entry point returns 0 on success and a nonzero MHD_ERR_* code on
 *    failure; mhd_last_error() then describes it.  The reference prints and
 *    STOPs instead (e.g. 3dmhdsub.f:1262-1265, 2942-2946); the shim does that.
 *  - One HIP stream per handle, one host thread per handle.
 *  - The whole domain is resident on the current HIP device.  The
 *    reference's MPI decomposition (3dmhd.f:87-88, 3dmhdsub.f:3031-3141)
 *    collapses to periodic index wrap in x/y and untouched z ghost planes.
 */
#ifndef RAST3DMHD_H
#define RAST3DMHD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MHD_OK = 0,
    MHD_ERR_ARG = 1,        /* bad argument or size */
    MHD_ERR_SCOPE = 2,      /* configuration outside the supported branches */
    MHD_ERR_HIP = 3,        /* HIP runtime error */
    MHD_ERR_STATE = 4,      /* call out of order (e.g. step before upload) */
    MHD_ERR_NONFINITE = 5   /* DT or UMACH not finite after a step */
};

/* Everything STEP reads from COMMON besides the arrays.  Replaces the
 * PARAMETER block (3dmhdparam.f: NPX,NPY,NPZ, LMAG, LROT, IZCON, ITCON,
 * IBCON) and the scalars of /GRID/, /RUNGKU/, /CPAR/, /CROT/, /CMAG/,
 * /CPEN/ and /CPER/ (3dmhd.f:62-71).  The float32-folded constants (C13,
 * GAM1, GAMMA, ...) must be passed as the driver computed them, never
 * retyped. */
typedef struct mhd_desc {
    int npx, npy, npz;          /* interior points; z includes both boundary planes */
    int lmag, lrot;             /* LMAG, LROT */
    int izcon, itcon, ibcon;    /* bottom T (0 fixed, 1 flux), top T (0 fixed, 1 flux, 2-4 below), B (0: BZ=0) */
    int precision;              /* MHD_FP64 (bit-identical to the reference) or MHD_MIXED */
    double hx, h2x, hy, h2y, hz, h2z, c13, c23, c43;   /* /GRID/ */
    double gam1, gam2, gam3, zeta1, zeta2;              /* /RUNGKU/ */
    double cv, ocv, ore, re, repr, theta, grav, sf, gamma; /* /CPAR/ */
    double omx, omz;                                    /* /CROT/ */
    double orm, rm, obeta;                              /* /CMAG/ */
    double tb;                                          /* /CPEN/ TB */
    double tp;                                          /* /CPER/ TP */
    double dzf;             /* float32 1.0E00/FLOAT(NPZ-1) widened (BCON, sub:2856/2929) */
    /* Optional branches of FLUXES and BCON (MHD_FP64 only; all zero: the
     * main path).  itcon may be 0..4: 2 embedded heat loss (sub:1783-1806),
     * 3 top T left to the equations, 4 TP cooler on half the top (2828-2850);
     * tp != 0 with itcon 0/1 is the plume boundary (2790-2827). */
    int id;                 /* 3dmhdparam.f ID: density diffusion 0, 1 (1/ROM), 2 (boundary-confined), sub:1301-1371 */
    int lrem, lshr;         /* LREM: Rempel conduction (sub:1536-1581); LSHR: shear boundaries (1493, 2751-2770) */
    int npey, npez;         /* the reference decomposition whose summation order the horizontal means
                               (HORIZONTAL_MEAN, sub:3143-3223) and the relaxation integral follow; 0 = 1 */
    double dh, dv;          /* 3dmhdparam.f DH, DV */
    double tu, dztu, dztb;  /* /SPECIALBOUND/ TU, DZTU, DZTB (LREM boundary values) */
    double rlax, tstart, toff;          /* /RELAX/ flux relaxation (sub:1615-1780), fixed-T top, fixed-flux bottom */
    double xp, yp, zp, tc, qfh, hh;     /* /CPER/ plume and heat-loss position, times and width */
    double zmax;            /* /BOUNDS/ ZMAX (ID=2) */
} mhd_desc;

/* Precision of the step (mhd_desc.precision). */
enum {
    MHD_FP64 = 0,   /* plain FP64, bit-identical to the reference build */
    MHD_MIXED = 1   /* FP32 perturbations about an FP64 hydrostatic background,
                       FP64 RK3 accumulation, boundary planes and dt reduction
                       (no counterpart in the reference) */
};

typedef struct mhd_ctx mhd_t;

/* Allocates the device state for one box.  Replaces the compile-time sizing
 * of COMMON /BIG/ (3dmhd.f:22-56). */
int mhd_create(const mhd_desc *desc, mhd_t **out);

/* Uploads the metric arrays of /AJACOBI/ and RKAPA/DKAPA of /CPEN/
 * (3dmhd.f:58-59, 65): lengths NX, NY, NZ (padded). */
int mhd_set_metrics(mhd_t *h,
                    const double *dxxdx, const double *d2xxdx2, const double *ddx,
                    const double *dyydy, const double *d2yydy2, const double *ddy,
                    const double *dzzdz, const double *d2zzdz2, const double *ddz,
                    const double *rkapa, const double *dkapa);

/* The coordinates EXX, WYY, ZEE of /AJACOBI/ (3dmhd.f:58-59; lengths NX,
 * NY, NZ).  Needed before the first step by the branches that use them: the
 * plume boundary (tp != 0, itcon 0/1), the embedded heat loss (itcon 2),
 * density diffusion ID=2 and the flux relaxation. */
int mhd_set_coordinates(mhd_t *h, const double *exx, const double *wyy, const double *zee);

/* MHD_MIXED only (no counterpart in the reference): the hydrostatic
 * background rho0(z), T0(z) about which perturbations are carried, padded
 * length NZ.  Call after mhd_set_metrics and before mhd_upload. */
int mhd_set_background(mhd_t *h, const double *ro0, const double *tt0);

/* MHD_MIXED only (no counterpart in the reference): re-centres the
 * background of the interior planes on the current horizontal means of RO
 * and TT and re-expresses the stored state about it, in place; the result
 * is what mhd_download, mhd_set_background(means) and mhd_upload would
 * give.  A background that has drifted from the flow's means leaves a
 * large horizontally uniform part in the stored pressure perturbation, and
 * its FP32 rounding then weighs on the horizontal pressure gradient.
 * mhd_set_rebase_interval(h, n) does it after every n-th step of mhd_step
 * and mhd_steps (0: never, the default); mhd_time_steps never does. */
int mhd_rebase(mhd_t *h);
int mhd_set_rebase_interval(mhd_t *h, int nsteps);

/* Uploads the prognostic arrays RU,RV,RW,RO,TT[,BX,BY,BZ] of /BIG/.  The
 * x/y ghosts are ignored (the library wraps periodically, which is what
 * COMMUNICATE establishes, 3dmhdsub.f:3131-3138); the z ghost planes are
 * kept.  bx, by, bz may be NULL when LMAG is off. */
int mhd_upload(mhd_t *h, const double *ru, const double *rv, const double *rw,
               const double *ro, const double *tt,
               const double *bx, const double *by, const double *bz);

/* Downloads the prognostic arrays, ghosts filled as COMMUNICATE leaves them
 * after STEP.  bx, by, bz may be NULL. */
int mhd_download(mhd_t *h, double *ru, double *rv, double *rw, double *ro, double *tt,
                 double *bx, double *by, double *bz);

/* One full time step: CFL/Mach reduction, three Wray RK3 stages with BCON
 * and the periodic ghost semantics (3dmhdsub.f:741-1173).  Returns the DT
 * used and the maximum Mach number UMACH computed from the state at entry
 * (COMMON /CTIM/ DT, /TRACE/ UMACH).  Synchronous. */
int mhd_step(mhd_t *h, double *dt, double *umach);

/* n steps back to back with no host synchronisation in between (the
 * driver's DO 1000 loop without dumps); dt/umach of the last step.  (With a
 * time-dependent plume boundary or heat-loss ramp, TC != 0, each step first
 * reads TIMT back: the host evaluates EXP/TANH of it as the reference's
 * Fortran runtime does.) */
int mhd_steps(mhd_t *h, int n, double *dt, double *umach);

/* Step counter and clocks of /ITER/ NIT and /CTIM/ TIMC, TIMT, TIMI as STEP
 * updates them (3dmhdsub.f:1171-1173). */
int mhd_get_clock(mhd_t *h, int *nit, double *timc, double *timt);
int mhd_set_clock(mhd_t *h, int nit, double timc, double timi);

/* No counterpart in the reference (which always steps with the CFL DT,
 * 3dmhdsub.f:827-871): dt > 0 makes every following step use this DT (the
 * CFL/Mach reduction still runs; UMACH and the non-finite check are
 * unchanged); dt = 0 returns to the CFL DT.  For runs at a fixed step and
 * for comparing forms at equal simulated times. */
int mhd_set_fixed_dt(mhd_t *h, double dt);

/* One evaluation of the right-hand side on the uploaded state, without
 * stepping: FU,FV,FW,FR,FT and (LMAG) WW1,WW2,WW3 = dB/dt, as FLUXES
 * leaves them (3dmhdsub.f:1178-2637).  Interior values only; ghost entries
 * of the output arrays are set to zero.  ww1..ww3 may be NULL. */
int mhd_fluxes(mhd_t *h, double *fu, double *fv, double *fw, double *fr, double *ft,
               double *ww1, double *ww2, double *ww3);

/* No counterpart in the reference (training data of the subgrid
 * closure): box filter of the state (source 0: RU..BZ) or of the
 * right-hand side at the state (source 1: FU,FV,FW,FR,FT,WW1..WW3) onto a
 * grid coarser by `factor`.  Horizontally the mean of factor x factor
 * blocks; vertically the mean of factor planes (zmode 0, NPZ divisible by
 * factor) or, for vertex grids with NPZ = factor*(nzc-1)+1 and even factor,
 * trapezoid weights over the factor spacings centred on every factor-th
 * plane, the two boundary planes taken as they are (zmode 1).  out is
 * [field][K][J][I] with I fastest, interior points only.  MHD_FP64 only. */
int mhd_box_filter(mhd_t *h, int factor, int zmode, int source, double *out);

/* Horizontal means and mean flux profiles of the current state (a
 * diagnostic, either precision), for the NPZ planes K=2..NPZ+1 (array
 * element k-1 for plane k = K-1):
 *   ttm, rom  HORIZONTAL_MEAN of TT and RO (3dmhdsub.f:3143-3223, the plain
 *             mean over the NPX*NPY interior points);
 *   fconm     mean convective flux, the horizontal mean of
 *             -RW*(2.5*(TT-TTM) + 0.5*|RU,RV,RW|^2/RO^2) (3dmhdsub.f:1650-1661);
 *   fradm     mean radiative flux (TTM(K+1)-TTM(K-1))*HZ*DZZDZ*RKAPA/REPR,
 *             one-sided on the two boundary planes (3dmhdsub.f:1663-1677).
 * The reference forms these only inside its flux-relaxation branch; here
 * they can be taken at any time.  The sums run in a fixed parallel order
 * (reproducible, equal to the reference's sequential sums to rounding).
 * Any pointer may be NULL; needs NPZ >= 3. */
int mhd_flux_profiles(mhd_t *h, double *ttm, double *rom, double *fconm, double *fradm);

/* Measurement hook (no counterpart in the reference): runs n steps like
 * mhd_steps with HIP events around every launch on the handle's stream and
 * returns the mean duration in milliseconds of one step, one RK stage
 * kernel launch (averaged over the three stages), one CFL reduction and one
 * boundary-plane launch.  Any output pointer may be NULL. */
int mhd_time_steps(mhd_t *h, int n, double *ms_step, double *ms_stage, double *ms_cfl,
                   double *ms_bcon);

/* Device pointers to the 11 metric arrays of mhd_set_metrics (same padded
 * lengths NX, NY, NZ), for mhd_step_noalloc. */
typedef struct mhd_dev_metrics {
    const double *dxxdx, *d2xxdx2, *ddx;
    const double *dyydy, *d2yydy2, *ddy;
    const double *dzzdz, *d2zzdz2, *ddz;
    const double *rkapa, *dkapa;
} mhd_dev_metrics;

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): one plain FP64 step, exactly as mhd_upload + mhd_step +
 * mhd_download on a handle, on arrays the caller keeps on the device.  No
 * handle, and no hipMalloc, hipFree or other allocation in the call.
 *  - desc: as for mhd_create; precision must be MHD_FP64.
 *  - met: device metric arrays.  state[0..7]: device pointers to the padded
 *    Fortran-order RU,RV,RW,RO,TT,BX,BY,BZ (state[5..7] may be NULL when
 *    LMAG is off), read at entry and overwritten with the stepped state,
 *    ghosts filled as mhd_download leaves them.
 *  - nit, timc, timt: /ITER/ and /CTIM/ clock in and out (NIT=NIT+1,
 *    TIMC=TIMC+DT, TIMT=TIMI+TIMC); timi is TIMI.  dt, umach: the DT used
 *    and UMACH of the state at entry (either may be NULL).
 *  - stream: a hipStream_t, or NULL for the null stream.  The call returns
 *    after the stream has finished the step.
 * The two RK registers, the Z register and the reduction scalars live in
 * one static device array of at most 64 MiB per GPU (held until the process
 * ends), so 3 * nf * 8 * (NPX*NPY*(NPZ+2) + 1184) bytes (nf = 8 with LMAG,
 * else 5; 1184 values of padding per field) must fit in 64 MiB - 256 B
 * (e.g. 64x64x32 MHD); larger boxes are rejected
 * with MHD_ERR_ARG.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other threads or streams) race.  Errors are reported
 * through mhd_last_error(NULL). */
int mhd_step_noalloc(const mhd_desc *desc, const mhd_dev_metrics *met, double *const *state,
                     int *nit, double *timc, double *timt, double timi,
                     double *dt, double *umach, void *stream);

/* Human-readable description of the last error on this handle (or of the
 * last failed mhd_create when h is NULL). */
const char *mhd_last_error(const mhd_t *h);

/* Frees the device state. */
void mhd_destroy(mhd_t *h);

#ifdef __cplusplus
}
#endif
#endif
