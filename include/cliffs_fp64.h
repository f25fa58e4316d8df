/* FP64 line mode of the cliffs_gpu library: the reference's state (celerity
 * c, xvel, yvel) carried in double precision on the device, each line swept
 * by the reference's own line procedure (cliffs.f:17-346: segment scan with
 * run-up expansion, Riemann invariants, ghosts, VTCS-2 integration,
 * limiter, edges) in the same per-node operation order, one wave of 64
 * lanes per line.  The segment scan, which the reference runs node after
 * node, is restated as per-node rules over a node's neighbours and a
 * prefix count across the wave, with the same outcome.  Its results equal
 * the reference's FP64 arithmetic bit for bit, for flow that amplifies
 * rounding (frictionless thin films on steep coasts, grids at the scheme's
 * stability limit) where the FP32 perturbation state of cliffs_gpu.h and
 * FP64 part.  It is a fidelity mode: a line is walked in 64-node chunks
 * through per-line work arrays, so a step costs far more than the FP32
 * fused step.
 *
 * Plain C ABI: pointers and sizes, column-major (Fortran) arrays as in
 * cliffs_gpu.h.  Return codes are those of cliffs_gpu.h (CLIFFS_OK, ...);
 * cliffs_last_error() describes the last failure.
 */
#ifndef CLIFFS_FP64_H
#define CLIFFS_FP64_H

#include "cliffs_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cliffs64_handle cliffs64_handle;

/* A grid in FP64 line mode (the same descriptor as cliffs_create; metrics
 * as the reference builds them, cliffs_main.f:195-235).  Maxima start at 0
 * (cliffs_main.f:238-239). */
int cliffs64_create(const cliffs_grid_desc *desc, cliffs64_handle **out);
int cliffs64_destroy(cliffs64_handle *h);

/* State in / out: cel (celerity sqrt(g H), celmin on dry land), xvel,
 * yvel, each (nXn, nYn) column-major. */
int cliffs64_upload_state(cliffs64_handle *h, const double *cel, const double *xvel,
                          const double *yvel);
int cliffs64_download_state(cliffs64_handle *h, double *cel, double *xvel, double *yvel);
/* hmax = max over steps of c, velmax = max of u^2 + v^2 (cliffs_main.f:347-355). */
int cliffs64_download_max(cliffs64_handle *h, double *hmax, double *velmax);

/* cliffs(iu, irw, n) for irw = first..last (1-based, inclusive): one
 * direction's sweep of those lines (cliffs_main.f:325-344).  edge1/edge2:
 * the reference's edge arrays (mx, 3) column-major -- normal velocity,
 * tangential velocity, elevation at the first and last node of each line
 * -- or NULL for an open edge at rest (zeros). */
int cliffs64_sweep(cliffs64_handle *h, int iu, int first, int last, const double *edge1,
                   const double *edge2, int mx);
/* The running maxima update of one step (cliffs_main.f:347-355). */
int cliffs64_maxwave(cliffs64_handle *h);
/* Gauge sampling (output_gages.f:18-29), as cliffs_gauges: eta = c^2/g - d,
 * or -1e32 where c <= celmin; u = xvel, v = yvel; I, J 1-based. */
int cliffs64_gauges(cliffs64_handle *h, int n, const int *I, const int *J, double *eta, double *u,
                    double *v);
/* Health of the current state, as cliffs_health: the largest Courant number
 * (|u| + c) dt / dx over wet cells and the number of non-finite cells. */
int cliffs64_health(cliffs64_handle *h, double *courant, long long *nonfinite);
/* One-way nesting from an FP64 parent, as cliffs_nest_create / _feed /
 * _destroy (the maps from cliffs_nest_maps): feed_children_grid
 * (feed_children_grid.f:10-179) on the parent's FP64 state, in the
 * reference's arithmetic -- the arrival test over the child footprint (|eta|
 * > 0.1*cuke or (|u|+|v|) c > cuke on cells with dep > hmin, latched in
 * *iarv), then (u, v, eta) blended onto the four child edges, dry parent
 * cells contributing 0; uvq_w/uvq_e (nyc,3), uvq_s/uvq_n (nxc,3) column-major;
 * *fed = 1 when the edges were filled.  A nest must be destroyed before its
 * parent. */
typedef struct cliffs64_nest cliffs64_nest;
int cliffs64_nest_create(cliffs64_handle *parent, int nxc, int nyc, const int *we_ny,
                         const double *we_cy, const int *sn_nx, const double *sn_cx, int w_nx,
                         double w_cx, int e_nx, double e_cx, int s_ny, double s_cy, int n_ny,
                         double n_cy, cliffs64_nest **out);
int cliffs64_nest_feed(cliffs64_nest *n, double cuke, double hmin, int *iarv, int *fed,
                       double *uvq_w, double *uvq_e, double *uvq_s, double *uvq_n);
int cliffs64_nest_destroy(cliffs64_nest *n);
/* nsteps full steps with open edges at rest: x-sweep of rows jj1..jj2,
 * y-sweep of columns ii1..ii2 (cliffs_main.f:140-164, 317-344), maxima. */
int cliffs64_run(cliffs64_handle *h, int nsteps);

#ifdef __cplusplus
}
#endif

#endif
