/*
 * rast3dmhd_ensemble.h -- batched ensemble of independent boxes.
 *
 * No counterpart in the reference, which steps one domain per MPI job.
 * An ensemble holds nbox independent boxes, each with its own sizes,
 * parameters (mhd_desc), metrics, state, adaptive DT and clock, resident on
 * the current HIP device.  One RK3 step of the whole ensemble is seven
 * launches: three persistent stage launches over a work table that covers
 * every box, three boundary-plane launches and one per-box DT/clock update.
 * Repeated steps are replayed from a captured HIP graph.  The tile shape,
 * z-chunk depth and persistent grid size come from mhd_ens_autotune,
 * mhd_ens_set_tiling or mhd_ens_set_tile.
 *
 * Every box advances exactly as the single-box library (rast3dmhd.h) would
 * advance it alone: the same kernel code per point, and reductions that
 * are exact min/max, so results are bit-identical whatever the tiling.
 *
 * Conventions are those of rast3dmhd.h: host FP64 arrays in the padded
 * Fortran order of the box, 0 on success, MHD_ERR_* on failure.
 */
#ifndef RAST3DMHD_ENSEMBLE_H
#define RAST3DMHD_ENSEMBLE_H

#include "rast3dmhd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhd_ens mhd_ens_t;

/* nbox boxes described by descs[0..nbox-1].  All boxes must share lmag and
 * precision (MHD_FP64 or MHD_MIXED). */
int mhd_ens_create(const mhd_desc *descs, int nbox, mhd_ens_t **out);

/* Per-box set-up, as mhd_set_metrics / mhd_set_background / mhd_upload /
 * mhd_download / mhd_set_clock / mhd_get_clock for box `box`. */
int mhd_ens_set_metrics(mhd_ens_t *e, int box,
                        const double *dxxdx, const double *d2xxdx2, const double *ddx,
                        const double *dyydy, const double *d2yydy2, const double *ddy,
                        const double *dzzdz, const double *d2zzdz2, const double *ddz,
                        const double *rkapa, const double *dkapa);
int mhd_ens_set_background(mhd_ens_t *e, int box, const double *ro0, const double *tt0);
int mhd_ens_upload(mhd_ens_t *e, int box, const double *ru, const double *rv, const double *rw,
                   const double *ro, const double *tt,
                   const double *bx, const double *by, const double *bz);
int mhd_ens_download(mhd_ens_t *e, int box, double *ru, double *rv, double *rw, double *ro, double *tt,
                     double *bx, double *by, double *bz);
int mhd_ens_set_clock(mhd_ens_t *e, int box, int nit, double timc, double timi);
int mhd_ens_get_clock(mhd_ens_t *e, int box, int *nit, double *timc, double *timt);
/* As mhd_set_fixed_dt for box `box`. */
int mhd_ens_set_fixed_dt(mhd_ens_t *e, int box, double dt);

/* n steps of every box (each with its own DT).  dt[b] and umach[b] receive
 * the last step's values of box b; either may be NULL.  use_graph selects
 * replay from a captured HIP graph (1) or direct launches (0). */
int mhd_ens_steps(mhd_ens_t *e, int n, int use_graph, double *dt, double *umach);

/* Tiling of the batched launches: tile_y (4 or 8 rows of 64 columns),
 * kdepth (planes per z chunk; 0 = whole column) and grid (persistent
 * workgroups; 0 = one workgroup per work item).  get_tiling reports the
 * rows of the current tile whatever its width. */
int mhd_ens_set_tiling(mhd_ens_t *e, int tile_y, int kdepth, int grid);
int mhd_ens_get_tiling(const mhd_ens_t *e, int *tile_y, int *kdepth, int *grid);

/* The same with the tile width: tile_x x tile_y is 64x4, 64x8, or (mixed
 * form only) 32x8, which leaves fewer lanes idle on boxes whose rows are
 * not a multiple of 64 points.  The results do not depend on the tiling. */
int mhd_ens_set_tile(mhd_ens_t *e, int tile_x, int tile_y, int kdepth, int grid);
int mhd_ens_get_tile(const mhd_ens_t *e, int *tile_x, int *tile_y, int *kdepth, int *grid);

/* Times every candidate tiling over nsteps steps (HIP events, direct
 * launches), keeps the fastest, and restores the state and clocks of every
 * box to what they were before the call.  report (may be NULL) receives a
 * JSON list of {tile_x, tile_y, kdepth, grid, ms_step} for the candidates
 * (tile shapes as in mhd_ens_set_tile; 7 chunk depths x 4 grids per shape,
 * about 80 bytes each: 8 KiB holds the list, a shorter report_len cuts it
 * at report_len - 1 bytes). */
int mhd_ens_autotune(mhd_ens_t *e, int nsteps, char *report, int report_len);

/* Mean milliseconds per ensemble step over n steps (HIP events on the
 * ensemble's stream, graph replay as in mhd_ens_steps with use_graph=1). */
int mhd_ens_time_steps(mhd_ens_t *e, int n, double *ms_step);

/* Zones of all boxes (interior points), for throughput figures. */
long long mhd_ens_zones(const mhd_ens_t *e);

const char *mhd_ens_last_error(const mhd_ens_t *e);
void mhd_ens_destroy(mhd_ens_t *e);

#ifdef __cplusplus
}
#endif
#endif
