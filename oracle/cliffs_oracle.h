/*
 * CPU oracle for the Cliffs VTCS-2 hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of the reference solver:
 *   cliffs(iu,irw,n)            cliffs.f:17-346
 *   sweep loops + max-wave      cliffs_main.f:317-355
 *   H -> celerity at init       cliffs_main.f:218-227
 *   grid metrics                MasterGrid.f:85-116
 *   pinpoint + nest maps        NestedGrid.f:98-148
 *   feed_children_grid          feed_children_grid.f:10-179
 * with the reference's operation order so that, compiled with
 * -ffp-contract=off, it is bit-exact to the reference (checked in tests/
 * against the reference's own cliffs.f built by oracle/Makefile).
 *
 * Only tests/, the cpu_baseline leg of bench.py and __graft_entry__.smoke()
 * may load this library, and only as the checker.  The product path never
 * calls it.
 *
 * Layout: every 2-D array is Fortran column-major (nx, ny), i.e. C
 * [ny][nx] with x fastest: element (i,j) (1-based) is a[(i-1) + nx*(j-1)].
 */
#ifndef CLIFFS_ORACLE_H
#define CLIFFS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* constants as the reference's FP32 literals widen them (global_modules.f:17) */
double orc_grav(void);
double orc_pi(void);

typedef struct orc_grid {
    int nx, ny;
    double *dep;          /* (nx,ny) */
    double *cel, *xvel, *yvel;   /* (nx,ny) state: celerity, x- and y-velocity */
    double *hmax, *velmax;       /* (nx,ny) */
    double *s1;           /* (nx-1,ny) */
    double *s2;           /* (ny-1) */
    double *zeta;         /* (ny) */
    double *edge1, *edge2;/* (mx,3) */
    int mx;
    double dt, ground, crough, celmin;
} orc_grid;

/* One line, one direction (cliffs.f:17-346).  iu=1: row irw along x;
 * iu=2: column irw along y.  1-based irw.  Returns the number of wet
 * segments found (kseg), or -1 if kseg exceeded the reference's fixed
 * 99-segment ghost storage (cliffs.f:26), where the reference's behaviour
 * is undefined.  The line is updated in either case: this restatement's
 * ghost storage grows with the line, as the GPU path has no segment limit. */
int orc_cliffs_line(orc_grid *g, int iu, int irw);

/* Sweep lines first..last (1-based, inclusive), OpenMP over lines
 * (cliffs_main.f:325-329, 340-344).  Returns the number of lines that hit
 * the 99-segment limit (0 in a valid run). */
int orc_sweep(orc_grid *g, int iu, int first, int last);

/* Max-wave update over all cells (cliffs_main.f:347-355). */
void orc_maxwave(orc_grid *g);

/* nsteps full steps with the grid's fixed edge arrays
 * (cliffs_main.f:317-355 without boundary-feed records). */
int orc_steps(orc_grid *g, int nsteps);

/* Line ranges of the time loop (cliffs_main.f:140-164). */
void orc_line_ranges(int nx, int ny, int *ii1, int *ii2, int *jj1, int *jj2,
                     int *xrun, int *yrun, int *mx);

/* H -> celerity (cliffs_main.f:218-226): c = 0 if H <= ground else sqrt(g H). */
void orc_depth_to_celerity(int n, const double *H, double ground, double *cel);

/* Grid metrics (MasterGrid.f:85-116).  s1 (nx-1,ny) if nx>2; s2 (ny-1) and
 * zeta (ny) if ny>2.  cartesian != 0 selects plain coordinate differences. */
void orc_metrics(int nx, int ny, const double *xcrd, const double *ycrd,
                 int cartesian, double *s1, double *s2, double *zeta);

/* pinpoint (NestedGrid.f:129-148): 1-based ip with the bisection the
 * reference uses, and weight cp. */
void orc_pinpoint(const double *xx, int nx, double xp, int *ip, double *cp);

/* Nest maps for one child (NestedGrid.f:104-119).  Outputs are 1-based
 * parent indices.  sn_nx/sn_cx have nxc entries, we_ny/we_cy have nyc. */
void orc_nest_maps(int nxp, const double *xp, int nyp, const double *yp,
                   int nxc, const double *xc, int nyc, const double *yc,
                   int *we_ny, double *we_cy, int *sn_nx, double *sn_cx,
                   int *w_nx, double *w_cx, int *e_nx, double *e_cx,
                   int *s_ny, double *s_cy, int *n_ny, double *n_cy);

/* feed_children_grid for one child (feed_children_grid.f:10-179).
 * *iarv is the arrival latch (in/out).  When arrived, fills
 * uvq_w/uvq_e (nyc,3) and uvq_s/uvq_n (nxc,3), column-major, with
 * (u, v, eta) and returns 1; returns 0 when the wave has not arrived. */
int orc_feed_child(const orc_grid *g, double cuke, double hmin,
                   int nxc, int nyc,
                   const int *we_ny, const double *we_cy,
                   const int *sn_nx, const double *sn_cx,
                   int w_nx, double w_cx, int e_nx, double e_cx,
                   int s_ny, double s_cy, int n_ny, double n_cy,
                   int *iarv,
                   double *uvq_w, double *uvq_e, double *uvq_s, double *uvq_n);

/* apply_initial_conditions (apply_initial_conditions.f:7-70): linear
 * (1-D) or bilinear (2-D) interpolation of qq(nx,ny) on lattice xq, yq,
 * ADDED onto pp(kx,ky) on lattice xp, yp. */
void orc_apply_initial_conditions(int nx, int ny, const double *xq, const double *yq,
                                  const double *qq, int kx, int ky, const double *xp,
                                  const double *yp, double *pp);

#ifdef __cplusplus
}
#endif
#endif
