/*
 * cliffs_gpu -- MI355X (gfx950) implementation of the Cliffs VTCS-2
 * dimensionally-split shallow-water step behind a plain C ABI.
 *
 * The reference solver keeps all state in Fortran modules and advances one
 * grid line per call:
 *
 *     subroutine cliffs(iu,irw,n)                   cliffs.f:17
 *
 * called from two OpenMP loops per time step (cliffs_main.f:325-329 rows,
 * 340-344 columns), followed by the max-wave update (cliffs_main.f:347-355).
 * This library owns the same state on the GPU (uploaded once, resident in
 * HBM) and replaces each whole OpenMP loop with one kernel launch.  The
 * Fortran side binds these entry points with ISO_C_BINDING (see the
 * package's fortran/ directory).
 *
 * Conventions
 *   - Every 2-D array is the reference's Fortran (nx, ny) column-major
 *     layout: element (i,j) (1-based) at [(i-1) + nx*(j-1)].
 *   - Line indices (irw, first, last) and gauge indices are 1-based, as in
 *     the reference.
 *   - Edge arrays are the reference's edge1/edge2(mx,3) (global_modules.f:34),
 *     column-major: (:,1) x-velocity, (:,2) y-velocity, (:,3) surface
 *     elevation.  NULL means all-zero edges (open boundary, still-water
 *     inflow; cliffs_main.f:288-290).
 *   - Every function returns CLIFFS_OK (0) or a negative error code and
 *     never aborts the process; cliffs_last_error() describes the failure.
 *   - A handle belongs to one host thread.  Compute calls are asynchronous
 *     on the handle's HIP stream; downloads synchronise.
 */
#ifndef CLIFFS_GPU_H
#define CLIFFS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define CLIFFS_OK 0
#define CLIFFS_ERR_ARG (-1)      /* bad argument (NULL, size, index range) */
#define CLIFFS_ERR_HIP (-2)      /* HIP runtime or launch failure */
#define CLIFFS_ERR_NOMEM (-3)    /* device or host allocation failed */
#define CLIFFS_ERR_METRIC (-4)   /* s1(i,j) is not a(j)*b(i) (MasterGrid.f:86-99 always is) */
#define CLIFFS_ERR_NODEVICE (-5) /* no HIP device / device index out of range */
#define CLIFFS_ERR_STATE (-6)    /* call out of order (e.g. compute before upload) */

typedef struct cliffs_handle cliffs_handle;

/* Grid and run parameters: the module data `cliffs` reads
 * (cliffs.f:19-20; global_modules.f:16-55). */
typedef struct cliffs_grid_desc {
    int nx, ny;          /* nXn, nYn */
    const double *dep;   /* dep(nXn,nYn): depth, positive = water */
    const double *s1;    /* s1(nXn-1,nYn): x spacing; may be NULL if nx <= 2 */
    const double *s2;    /* s2(nYn-1): y spacing; may be NULL if ny <= 2 */
    const double *zeta;  /* zeta(nYn): spherical / channel-width term; may be NULL if ny <= 2 */
    double dt;           /* time step (PARAMETERS dt) */
    double ground;       /* minimal flow depth, > 0 (PARAMETERS ground) */
    double crough;       /* Manning n^2 (PARAMETERS crough) */
    int device;          /* HIP device ordinal */
} cliffs_grid_desc;

/* Allocate device state for one grid.  Replaces the allocation of the
 * MASTER arrays (cliffs_main.f:191-197) and precomputes celmin
 * (cliffs_main.f:227) and the device form of dep/s1/s2/zeta. */
int cliffs_create(const cliffs_grid_desc *desc, cliffs_handle **out);
/* Frees the grid.  Fails with CLIFFS_ERR_STATE (and frees nothing) while a
 * nest (as parent or coupled child) or an ensemble still uses it. */
int cliffs_destroy(cliffs_handle *h);

/* cel (celerity sqrt(g H), cliffs_main.f:218-226), xvel, yvel: FP64
 * (nx,ny).  Upload converts to the device's FP32 perturbation form so that
 * the wet/dry masks of the next sweep equal the FP64 ones bit for bit. */
int cliffs_upload_state(cliffs_handle *h, const double *cel, const double *xvel,
                        const double *yvel);
int cliffs_download_state(cliffs_handle *h, double *cel, double *xvel, double *yvel);

/* Running maxima hmax = max c, velmax = max(u^2+v^2) (cliffs_main.f:347-355). */
int cliffs_download_max(cliffs_handle *h, double *hmax, double *velmax);
int cliffs_reset_max(cliffs_handle *h); /* hmax = velmax = 0 (cliffs_main.f:196-197) */

/* One directional sweep over lines first..last (1-based, inclusive):
 * iu=1 rows along x, iu=2 columns along y.  Equivalent to
 *     do irw = first, last; call cliffs(iu, irw, n); end do
 * (cliffs_main.f:325-329 / 340-344).  edge1/edge2 as described above, mx
 * their leading dimension. */
int cliffs_sweep(cliffs_handle *h, int iu, int first, int last, const double *edge1,
                 const double *edge2, int mx);

/* Max-wave update over all cells (cliffs_main.f:347-355). */
int cliffs_maxwave(cliffs_handle *h);

/* One full time step (cliffs_main.f:317-355): x-sweep over rows jj1..jj2,
 * y-sweep over columns ii1..ii2, max-wave -- with the 1-D/2-D line ranges
 * of cliffs_main.f:140-164.  xedge1/xedge2 are the W/E edges of the x-sweep,
 * yedge1/yedge2 the S/N edges of the y-sweep, already interpolated in time
 * by the caller (cliffs_main.f:318-324, 333-339). */
int cliffs_step(cliffs_handle *h, const double *xedge1, const double *xedge2,
                const double *yedge1, const double *yedge2, int mx);

/* nsteps steps with zero edges (open boundaries), enqueued without host
 * synchronisation. */
int cliffs_run(cliffs_handle *h, int nsteps);

/* Like cliffs_run, bracketing the kernel launches of every 8th step (the
 * first included) with HIP events on the handle's stream -- an event record
 * between launches costs the stream a few microseconds, so the other steps
 * run without.  Returns the device time from the first to the last launch
 * (ms_total) and mean kernel durations over the bracketed steps: of the x-
 * and y-sweep kernels, or, when the step runs fused (2-D grids, the
 * default), of the fused step kernel (ms_xsweep) and of its flag-check
 * kernel (ms_ysweep).  Events are reused in a ring, so any nsteps works;
 * the call returns when the work has finished. */
int cliffs_run_timed(cliffs_handle *h, int nsteps, double *ms_total, double *ms_xsweep,
                     double *ms_ysweep);

/* nsamples x `every` steps with open edges, recording the surface elevation
 * at n gauges (1-based I, J; eta = c^2/g - d on wet cells, 0 on dry ones)
 * after every `every` steps: eta[k*n + q] for sample k, gauge q (FP32).
 * Recorded on the device; one copy to the host at the end. */
int cliffs_run_gauges(cliffs_handle *h, int every, int nsamples, int n, const int *I, const int *J,
                      float *eta);
/* cliffs_run_gauges without any device allocation: the gauge table and the
 * series live in a library-wide device scratch of 2^20 4-byte words (4 MiB)
 * per GPU, held until the process ends; fails with CLIFFS_ERR_ARG when
 * n*(nsamples+2) exceeds it.  Results equal cliffs_run_gauges'.  Calls that
 * overlap on one GPU (handles used from several host threads) race on the
 * scratch.  No counterpart in the reference: for callers that cannot
 * allocate device memory (serving loops and harnesses that forbid
 * hipMalloc). */
int cliffs_run_gauges_noalloc(cliffs_handle *h, int every, int nsamples, int n, const int *I,
                              const int *J, float *eta);

/* 2-D steps (cliffs_step, cliffs_run) run as one fused kernel per step
 * (on = 1, the default) or as an x-sweep kernel followed by a y-sweep kernel
 * (on = 0).  Results are identical; the fused step moves 44 instead of 72
 * bytes of HBM traffic per cell.  No counterpart in the reference, whose
 * step is always two sweeps (cliffs_main.f:317-345). */
int cliffs_set_fused(cliffs_handle *h, int on);
/* Tile height of the fused step: at most rows = 8..256 (the grid's rows are
 * split into as many chunks as that height needs, of equal height), or 0
 * for the automatic choice (enough tiles to fill the device once, at least
 * 32 rows).  Results do not depend on it; small grids step faster with
 * short tiles (cliffs_ensemble_tune picks one by timing).  Fails with
 * CLIFFS_ERR_STATE for a member of an ensemble (its tuner owns the tile
 * height).  No counterpart in the reference. */
int cliffs_set_tile_rows(cliffs_handle *h, int rows);
/* Tile width of the fused step: 248 columns (a 256-node x window, four
 * nodes per lane; the default) or 120 (a 128-node window, two per lane),
 * with the automatic tile height.  Results do not depend on it.  Fails with
 * CLIFFS_ERR_STATE for a member of an ensemble.  No counterpart in the
 * reference. */
int cliffs_set_tile_width(cliffs_handle *h, int width);

/* ---- ensembles: many grids stepped together ---------------------------- */

typedef struct cliffs_ensemble cliffs_ensemble;

/* Group n grids (2-D, state uploaded, one device) to advance together: one
 * step of the ensemble is one time step (cliffs_main.f:317-355, open edges)
 * of every member, all members' tiles in one launch of the fused step
 * kernel, and consecutive steps are replayed from a captured HIP graph.
 * Each member's results are bit-identical to stepping it alone.  No
 * counterpart in the reference, which runs one scenario per process.  The
 * members stay owned by the caller and usable between ensemble calls (calls
 * are ordered across the members' streams and the ensemble's); destroy the
 * ensemble before any member. */
int cliffs_ensemble_create(cliffs_handle *const *members, int n, cliffs_ensemble **out);
int cliffs_ensemble_destroy(cliffs_ensemble *e);
int cliffs_ensemble_run(cliffs_ensemble *e, int nsteps);
/* As cliffs_ensemble_run; ms_total = device time of the nsteps. */
int cliffs_ensemble_run_timed(cliffs_ensemble *e, int nsteps, double *ms_total);
/* Autotuner: times trial_steps repetitions of one ensemble step for each
 * candidate tile shape (widths 248 and 120 columns, heights 16 .. 256 rows)
 * and keeps the fastest (its height in rows_out, both in
 * cliffs_ensemble_shape).  The trials change neither the state nor the
 * running maxima (their max-wave update is switched off). */
int cliffs_ensemble_tune(cliffs_ensemble *e, int trial_steps, int *rows_out);
/* The tile shape the ensemble steps with (after cliffs_ensemble_tune: the
 * fastest; rows 0 = each member's automatic height).  Members must share one
 * tile width when the ensemble is created. */
int cliffs_ensemble_shape(cliffs_ensemble *e, int *width, int *rows);
int cliffs_ensemble_sync(cliffs_ensemble *e);
/* nsamples x `every` steps of every member (`every` even); after each
 * `every` steps the surface elevation at each member's n gauges (1-based
 * I[m*n + q], J[m*n + q]; eta on wet cells, 0 on dry, as output_gages.f:18-29
 * writes it) goes to eta[(m*nsamples + k)*n + q].  One gauge launch per
 * sample for all members; the series stay on the device until the end. */
int cliffs_ensemble_run_gauges(cliffs_ensemble *e, int every, int nsamples, int n, const int *I,
                               const int *J, float *eta);

/* Gauge sampling (output_gages.f:18-29): eta = c^2/g - d, or -1e32 (the
 * reference's real*4 NAN) where c <= celmin; u = xvel, v = yvel.  I, J are
 * 1-based. */
int cliffs_gauges(cliffs_handle *h, int n, const int *I, const int *J, double *eta,
                  double *u, double *v);

/* ---- one-way nesting: parent -> child boundary feed -------------------- */

typedef struct cliffs_nest cliffs_nest;

/* Parent index maps of one child grid (NestedGrid.f:104-119), computed on the
 * host with the reference's bisection `pinpoint` (NestedGrid.f:129-148) in
 * FP64.  xp/yp: parent lattice (nXn, nYn); xc/yc: child lattice (nxc, nyc).
 * Outputs are 1-based parent indices and FP64 weights, laid out as the
 * NESTED module holds them for one child: we_ny/we_cy (nyc), sn_nx/sn_cx
 * (nxc), and the scalars of the four edges. */
int cliffs_nest_maps(int nxp, const double *xp, int nyp, const double *yp, int nxc,
                     const double *xc, int nyc, const double *yc, int *we_ny, double *we_cy,
                     int *sn_nx, double *sn_cx, int *w_nx, double *w_cx, int *e_nx,
                     double *e_cx, int *s_ny, double *s_cy, int *n_ny, double *n_cy);

/* Attach one child's maps (as produced by NestedGrid or cliffs_nest_maps)
 * to a parent handle; the maps are copied to the device. */
int cliffs_nest_create(cliffs_handle *parent, int nxc, int nyc, const int *we_ny,
                       const double *we_cy, const int *sn_nx, const double *sn_cx, int w_nx,
                       double w_cx, int e_nx, double e_cx, int s_ny, double s_cy, int n_ny,
                       double n_cy, cliffs_nest **out);
int cliffs_nest_destroy(cliffs_nest *n);

/* feed_children_grid (feed_children_grid.f:10-179) on the device state:
 * while *iarv == 0, test wave arrival over the child footprint (|eta| >
 * 0.1*cuke, or (|u|+|v|) c > cuke, on cells with dep > hmin) and latch *iarv;
 * once arrived, interpolate (u, v, eta) onto the four child edges, dry parent
 * cells contributing 0.  uvq_w/uvq_e are (nyc,3), uvq_s/uvq_n (nxc,3),
 * column-major -- the arrays output_feed writes (output_feed.f:3-24).
 * *fed is set to 1 when the edges were filled. */
int cliffs_nest_feed(cliffs_nest *n, double cuke, double hmin, int *iarv, int *fed,
                     double *uvq_w, double *uvq_e, double *uvq_s, double *uvq_n);

/* ---- one-way nesting on one device: parent and child stepped together -- */

/* Couple the nest to a child grid (nxc x nyc, same device, 2-D) so that the
 * parent's feed records stay in device memory -- a ring of `capacity`
 * records -- and the child's edges are interpolated in time on the device.
 * Replaces the file round trip feed_children_grid -> output_feed ->
 * get_boundary_feed -> edge interpolation (cliffs_main.f:304-339) of the
 * reference, which runs parent and child as separate processes. */
int cliffs_nest_couple(cliffs_nest *n, cliffs_handle *child, int capacity);
/* (Destroy a nest before its parent and child grids, and an ensemble before
 * its members: they use the grids' device memory and streams;
 * cliffs_destroy refuses a grid still in use.) */
/* feed_children_grid at parent time t: arrival test (the latch iarv as in
 * cliffs_nest_feed), then, once arrived, one more feed record (*nrec = records
 * so far).  Fails with CLIFFS_ERR_STATE if the record would overwrite one
 * the child still needs. */
int cliffs_nest_record(cliffs_nest *n, double cuke, double hmin, double t, int *iarv, int *nrec);
/* No more records will come: once past the last record the child's edges
 * become 0 (cliffs_main.f:306-310). */
int cliffs_nest_end_feed(cliffs_nest *n);
/* The child's start (get_boundary_feed.f:144-167, cliffs_main.f:188): the
 * first feed record whose largest |eta| over the four edges is >= cuke (the
 * child's own threshold, its parameter file's "still sea threshold"); *t0 =
 * that record's time, where the child's clock starts.  Fails with
 * CLIFFS_ERR_STATE while no record has reached cuke yet (from then on each
 * new record is checked as cliffs_nest_record writes it, and records below
 * the threshold leave the ring; call again for the start), when
 * the feed ended without one, or when only its last record reaches it (the
 * reference's "bndr signal below threshold").  Without this call the child
 * starts at the first record (the rule with cuke = 0). */
int cliffs_nest_child_start(cliffs_nest *n, double cuke, double *t0);
/* One child step at child time t (after t = t + dt): records irec-1, irec
 * with tinput(irec-1) <= t < tinput(irec), starting from the start record,
 * edges = rec(irec-1)*cc + rec(irec)*(1-cc), cc = (tinput(irec)-t) /
 * (tinput(irec)-tinput(irec-1)) (cliffs_main.f:305-339), then the step;
 * irec counts from the child's start record (cliffs_nest_child_start), and
 * t must not precede that record's time.
 * Fails with CLIFFS_ERR_STATE if t is beyond the parent's last record and
 * the feed has not ended. */
int cliffs_nest_child_step(cliffs_nest *n, double t);
/* A coupled run on one device: the loop a driver writes around cliffs_run,
 * cliffs_nest_record and cliffs_nest_child_step (cliffs_main.f:288-365 with
 * the child stepped alongside instead of from files), in one call.  For
 * each of `nrec` record intervals the parent advances `every` steps of dt,
 * a record is written at parent time t0 + k*every*dt (k = intervals this
 * nest has run, from 1; arrival test with cuke and hmin as in
 * cliffs_nest_record, latch *iarv), and the child advances, `per` steps of
 * dt_child per record interval, as far as the records written so far allow:
 * its step kc (from 1) at time tinput(s + kc/per) + (kc mod per)*dt_child,
 * s its start record.  The counts carry over between calls, so a run may
 * be split into any number of calls. */
int cliffs_nest_run(cliffs_nest *n, int nrec, int every, double t0, double dt, int per,
                    double dt_child, double cuke, double hmin, int *iarv);

/* Health of the current state: *courant = the largest Courant number
 * (|u| + c) dt / dx over wet cells (x along rows where the x-sweep runs, y
 * along columns where the y-sweep runs, dx the smaller spacing next to the
 * cell), *nonfinite = the number of cells whose c, u or v is NaN or
 * infinite.  One reduction over the state on the device; waits for it.  No
 * counterpart in the reference, which checks neither (cliffs_main.f
 * integrates with a fixed dt); for drivers that want to detect a blow-up. */
int cliffs_health(cliffs_handle *h, double *courant, long long *nonfinite);

int cliffs_sync(cliffs_handle *h);
/* The handle's hipStream_t, for callers that enqueue their own work. */
int cliffs_stream(cliffs_handle *h, void **stream);
/* Gravity, celmin and line ranges as the device uses them. */
int cliffs_info(cliffs_handle *h, double *grav, double *celmin, int *ii1, int *ii2, int *jj1,
                int *jj2);
const char *cliffs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CLIFFS_GPU_H */
