/*
 * CPU oracle for the explicit compressible-MHD step (TEST INFRASTRUCTURE ONLY).
 * See mhd_oracle.c for the scope and the reference lines each routine follows.
 */
#ifndef MHD_ORACLE_H
#define MHD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the reference COMMON blocks that STEP reads and writes
 * (3dmhd.f:51-71).  Arrays are padded Fortran-order (nx,ny,nz); metric
 * arrays have length nx / ny / nz. */
typedef struct orc_state {
    int nx, ny, nz;            /* padded sizes NPX+2, NPY+2, NPZ+2 */
    int lmag, lrot;            /* physics switches (3dmhdparam.f:8) */
    int izc, itc, ibc;         /* BC codes IZCON, ITCON, IBCON */
    int nit;                   /* /ITER/ NIT */
    /* /GRID/ */
    double hx, h2x, hy, h2y, hz, h2z, c13, c23, c43;
    /* /RUNGKU/ */
    double gam1, gam2, gam3, zeta1, zeta2;
    /* /CPAR/ */
    double cv, ocv, ore, re, repr, theta, grav, sf, gamma;
    /* /CROT/, /CMAG/, /CPEN/ TB, /CPER/ TP */
    double omx, omz, orm, rm, obeta, tb, tp;
    /* float32 1.0E00/FLOAT(NPZ-1) widened to double (BCON DZ, sub:2856/2929) */
    double dzf;
    /* /CTIM/, /TRACE/ */
    double dt, timt, timc, timi, umach;
    double cflmin[4];          /* the four CFL minima before SF*MIN */
    /* /AJACOBI/ and /CPEN/ metric arrays */
    const double *dxxdx, *d2xxdx2, *ddx;
    const double *dyydy, *d2yydy2, *ddy;
    const double *dzzdz, *d2zzdz2, *ddz;
    const double *rkapa, *dkapa;
    /* /BIG/ */
    double *ru, *rv, *rw, *ro, *tt;
    double *uu, *vv, *ww;
    double *fu, *fv, *fw, *fr, *ft;
    double *zru, *zrv, *zrw, *zro, *ztt;
    double *ww1, *ww2, *ww3;
    double *bx, *by, *bz;
    double *zbx, *zby, *zbz;
    /* Optional branches of FLUXES and BCON (all zero: the main path).
     *   id, dh, dv      density diffusion (3dmhdparam.f ID, DH, DV)
     *   lrem, tu, dztu, dztb   Rempel conduction, /SPECIALBOUND/
     *   lshr            shear boundaries
     *   rlax, tstart, toff     flux relaxation, /RELAX/
     *   xp..hh          /CPER/ (TP above): plume boundary, embedded heat loss
     *   zmax            /BOUNDS/ ZMAX (ID=2)
     *   npey, npez      the MPI decomposition whose summation order the
     *                   horizontal means and the relaxation integral follow
     *   exx, wyy, zee   /AJACOBI/ coordinates (lengths nx, ny, nz)
     *   icall, ttm_min, ttm_max, tfac   FLUXES' saved relaxation state
     *                   (DATA 0, 1D99, -1D99, 1D0) */
    int id, lrem, lshr, npey, npez, icall;
    double dh, dv, tu, dztu, dztb, rlax, tstart, toff;
    double xp, yp, zp, tc, qfh, hh, zmax;
    double ttm_min, ttm_max, tfac;
    const double *exx, *wyy, *zee;
} orc_state;

int orc_check(const orc_state *s);
void orc_fluxes(orc_state *s);
void orc_bcon(orc_state *s);
void orc_communicate(orc_state *s);
void orc_cfl(orc_state *s);
void orc_step(orc_state *s);
/* Horizontal means TTM, ROM and the mean convective / radiative flux
 * profiles FCONM, FRADM (diagnostic; arrays of NZ doubles, padded K). */
void orc_flux_profiles(const orc_state *s, double *ttm, double *rom, double *fconm, double *fradm);

#ifdef __cplusplus
}
#endif
#endif
