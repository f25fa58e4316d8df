/*
 * CPU oracle for the Cliffs VTCS-2 hot path -- TEST INFRASTRUCTURE ONLY.
 * See cliffs_oracle.h for scope.  Every expression keeps the reference's
 * evaluation order (Fortran left-to-right for equal precedence) so that
 * the result is bit-identical to cliffs.f built at -O0 without FMA.
 */
#include "cliffs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* global_modules.f:17 -- default-REAL (FP32) literals assigned to real*8 */
static const double GRAV = (double)9.8f;
static const double PI_R = (double)3.1415926536f;
static const double EARTHR = 6371009.0;
/* cliffs.f:34 -- FP32 literal */
static const double SQR2 = (double)1.41421356f;

double orc_grav(void) { return GRAV; }
double orc_pi(void) { return PI_R; }

/* per-thread scratch, 1-based arrays of length n+1 */
typedef struct {
    int cap;
    double *h, *pp, *qq, *vv, *dx, *scl, *depth, *Pinv, *Qinv, *v, *u;
    double *pqvdL, *pqvdR; /* 4 x (cap+1) */
    int *lghost, *rghost;
    unsigned char *land, *newland;
} orc_ws;

static void ws_free(orc_ws *w) {
    free(w->h); free(w->pp); free(w->qq); free(w->vv); free(w->dx);
    free(w->scl); free(w->depth); free(w->Pinv); free(w->Qinv); free(w->v);
    free(w->u); free(w->pqvdL); free(w->pqvdR); free(w->lghost);
    free(w->rghost); free(w->land); free(w->newland);
    memset(w, 0, sizeof(*w));
}

static int ws_reserve(orc_ws *w, int n) {
    if (w->cap >= n) return 0;
    ws_free(w);
    size_t m = (size_t)n + 2;
    w->h = malloc(m * sizeof(double)); w->pp = malloc(m * sizeof(double));
    w->qq = malloc(m * sizeof(double)); w->vv = malloc(m * sizeof(double));
    w->dx = malloc(m * sizeof(double)); w->scl = malloc(m * sizeof(double));
    w->depth = malloc(m * sizeof(double)); w->Pinv = malloc(m * sizeof(double));
    w->Qinv = malloc(m * sizeof(double)); w->v = malloc(m * sizeof(double));
    w->u = malloc(m * sizeof(double));
    w->pqvdL = malloc(4 * m * sizeof(double));
    w->pqvdR = malloc(4 * m * sizeof(double));
    w->lghost = malloc(m * sizeof(int)); w->rghost = malloc(m * sizeof(int));
    w->land = malloc(m); w->newland = malloc(m);
    if (!w->h || !w->pp || !w->qq || !w->vv || !w->dx || !w->scl || !w->depth ||
        !w->Pinv || !w->Qinv || !w->v || !w->u || !w->pqvdL || !w->pqvdR ||
        !w->lghost || !w->rghost || !w->land || !w->newland) {
        ws_free(w);
        return -1;
    }
    w->cap = n;
    return 0;
}

#define PQL(r, k) w->pqvdL[4 * (k) + (r) - 1]
#define PQR(r, k) w->pqvdR[4 * (k) + (r) - 1]

/* cliffs.f:17-346 */
static int cliffs_line_ws(orc_grid *g, int iu, int irw, orc_ws *w) {
    const int nx = g->nx;
    const int n = (iu == 1) ? g->nx : g->ny;
    const double celmin = g->celmin, ground = g->ground, dt = g->dt;
    const double crough = g->crough;
    double *h = w->h, *pp = w->pp, *qq = w->qq, *vv = w->vv, *dx = w->dx;
    double *scl = w->scl, *depth = w->depth, *Pinv = w->Pinv, *Qinv = w->Qinv;
    double *v = w->v, *u = w->u;
    int *lghost = w->lghost, *rghost = w->rghost;
    unsigned char *land = w->land, *newland = w->newland;
    int i, j, k, i1, i2, kseg, k1, k2, water;
    double flood;

    /* gather (cliffs.f:36-54) */
    if (iu == 1) {
        const size_t row = (size_t)nx * (size_t)(irw - 1);
        for (i = 1; i <= n; i++) {
            depth[i] = g->dep[row + i - 1];
            h[i] = g->cel[row + i - 1];
            u[i] = g->xvel[row + i - 1];
            v[i] = g->yvel[row + i - 1];
            scl[i] = 0;
        }
        for (i = 1; i <= n - 1; i++) dx[i] = g->s1[(size_t)(nx - 1) * (irw - 1) + i - 1];
    } else {
        for (i = 1; i <= n; i++) {
            const size_t o = (size_t)(irw - 1) + (size_t)nx * (i - 1);
            depth[i] = g->dep[o];
            h[i] = g->cel[o];
            u[i] = g->yvel[o];
            v[i] = g->xvel[o];
            scl[i] = g->zeta[i - 1];
        }
        for (i = 1; i <= n - 1; i++) dx[i] = g->s2[i - 1];
    }

    for (i = 1; i <= n; i++) { lghost[i] = 0; rghost[i] = 0; }
    /* dry areas (cliffs.f:58-63) */
    for (j = 1; j <= n; j++) land[j] = (h[j] <= celmin);
    for (j = 1; j <= n; j++) newland[j] = land[j];

    /* segment scan with run-up expansion (cliffs.f:65-111) */
    water = 0;
    kseg = 0;
    if (!land[1]) { water = 1; kseg = 1; lghost[1] = 1; }
    k1 = kseg + 1;
    for (i = 2; i <= n; i++) {
        if (water) {
            if (land[i]) {
                flood = (h[i - 1] * h[i - 1]) / GRAV - depth[i - 1] + depth[i];
                if ((flood > ground) && (!land[i - 1]) && (u[i - 1] >= 0)) {
                    u[i] = u[i - 1];
                    v[i] = v[i - 1];
                    h[i] = celmin;
                    newland[i] = 0;
                } else {
                    water = 0;
                    rghost[kseg] = i;
                }
            }
        } else if (!land[i]) {
            water = 1;
            kseg = kseg + 1;
            flood = (h[i] * h[i]) / GRAV - depth[i] + depth[i - 1];
            if ((flood > ground) && (u[i] <= 0)) {
                lghost[kseg] = (i - 2 > 1) ? i - 2 : 1;
                newland[i - 1] = 0;
                u[i - 1] = u[i];
                v[i - 1] = v[i];
                h[i - 1] = celmin;
            } else {
                lghost[kseg] = i - 1;
            }
            if (kseg > 1 && lghost[kseg] < rghost[kseg - 1]) {
                lghost[kseg] = 0;
                kseg = kseg - 1;
                rghost[kseg] = 0;
            }
        }
    }

    if (kseg == 0) return 0; /* cliffs.f:113 -- line left untouched */
    if (rghost[kseg] == 0) rghost[kseg] = n;
    k2 = kseg - 1;
    if (newland[rghost[kseg]]) k2 = kseg;

    for (i = 1; i <= n; i++) {
        if (newland[i]) { h[i] = 0; u[i] = 0; v[i] = 0; }
        vv[i] = v[i];
    }

    /* Riemann invariants (cliffs.f:125-136) */
    for (i = 1; i <= n; i++) { Pinv[i] = 0; Qinv[i] = 0; }
    for (k = 1; k <= kseg; k++) {
        i1 = lghost[k] + 1;
        i2 = rghost[k] - 1;
        for (i = i1; i <= i2; i++) {
            Pinv[i] = u[i] + 2 * h[i];
            Qinv[i] = u[i] - 2 * h[i];
        }
    }
    /* ghosts (cliffs.f:137-152) */
    if (k1 <= kseg) {
        for (k = k1; k <= kseg; k++) {
            j = lghost[k] + 1;
            PQL(1, k) = -Qinv[j]; PQL(2, k) = -Pinv[j]; PQL(3, k) = v[j]; PQL(4, k) = depth[j];
        }
    }
    if (k2 >= 1) {
        for (k = 1; k <= k2; k++) {
            j = rghost[k] - 1;
            PQR(1, k) = -Qinv[j]; PQR(2, k) = -Pinv[j]; PQR(3, k) = v[j]; PQR(4, k) = depth[j];
        }
    }
    /* real edge values (cliffs.f:153-164) */
    if (!newland[1]) {
        Pinv[1] = u[1] + 2 * h[1];
        Qinv[1] = u[1] - 2 * h[1];
        PQL(1, 1) = Pinv[1]; PQL(2, 1) = Qinv[1]; PQL(3, 1) = v[1]; PQL(4, 1) = depth[1];
    }
    if (!newland[n]) {
        Pinv[n] = u[n] + 2 * h[n];
        Qinv[n] = u[n] - 2 * h[n];
        PQR(1, kseg) = Pinv[n]; PQR(2, kseg) = Qinv[n]; PQR(3, kseg) = v[n]; PQR(4, kseg) = depth[n];
    }
    for (i = 1; i <= n; i++) { qq[i] = Qinv[i]; pp[i] = Pinv[i]; }

    /* time-step integration (cliffs.f:168-253) */
    for (k = 1; k <= kseg; k++) {
        i1 = lghost[k] + 1;
        i2 = rghost[k] - 1;
        if (i1 > i2) continue;
        for (i = i1; i <= i2; i++) {
            double p1, q1, v1, d1, p2, q2, v2, d2;
            double uj, vj, hj, cc, pj, qj, sphr, fm, fp, dtx, dj, dpj, dpm;
            double ej, e2, e1, Qpj, Qjm, Qsum;
            int wet3;
            if (i == i1) { p1 = PQL(1, k); q1 = PQL(2, k); v1 = PQL(3, k); d1 = PQL(4, k); }
            else { p1 = Pinv[i - 1]; q1 = Qinv[i - 1]; v1 = v[i - 1]; d1 = depth[i - 1]; }
            if (i == i2) { p2 = PQR(1, k); q2 = PQR(2, k); v2 = PQR(3, k); d2 = PQR(4, k); }
            else { p2 = Pinv[i + 1]; q2 = Qinv[i + 1]; v2 = v[i + 1]; d2 = depth[i + 1]; }
            uj = u[i];
            vj = v[i];
            hj = h[i] * h[i] / GRAV;
            /* almost-Manning friction (cliffs.f:204) */
            cc = GRAV * crough * uj * sqrt((uj * uj + vj * vj) / hj) / hj;

            pj = Pinv[i];
            qj = Qinv[i];
            sphr = (pj - qj) * (pj + qj) * scl[i];
            fm = dx[i - 1];
            fp = dx[i];
            dtx = dt / (fm + fp);
            dj = depth[i];
            dpj = GRAV * (d2 - dj);
            dpm = GRAV * (dj - d1);
            wet3 = (dj > ground) && (d1 > ground) && (d2 > ground);

            ej = 3 * pj + qj;
            e2 = 3 * p2 + q2;
            e1 = 3 * p1 + q1;
            Qpj = (0.125 * (e2 + ej) * (p2 - pj) - dpj) / fp;
            Qjm = (0.125 * (ej + e1) * (pj - p1) - dpm) / fm;
            if ((e1 < 0) && (e2 > 0) && wet3)
                Qsum = (0.25 * ej * (p2 - p1) - 2 * GRAV * (sqrt(dj * d2) - sqrt(dj * d1))) / (fm + fp);
            else
                Qsum = (Qpj + Qjm) / 2;
            Qsum = Qsum + 0.25 * ej * dtx * (Qjm - Qpj);
            pp[i] = pj - dt * (Qsum + cc - sphr);

            ej = 3 * qj + pj;
            e2 = 3 * q2 + p2;
            e1 = 3 * q1 + p1;
            Qpj = (0.125 * (e2 + ej) * (q2 - qj) - dpj) / fp;
            Qjm = (0.125 * (ej + e1) * (qj - q1) - dpm) / fm;
            if ((e1 < 0) && (e2 > 0) && wet3)
                Qsum = (0.25 * ej * (q2 - q1) - 2 * GRAV * (sqrt(dj * d2) - sqrt(dj * d1))) / (fm + fp);
            else
                Qsum = (Qpj + Qjm) / 2;
            Qsum = Qsum + 0.25 * ej * dtx * (Qjm - Qpj);
            qq[i] = qj - dt * (Qsum + cc + sphr);

            Qsum = (v2 - vj) / fp - (vj - v1) / fm;
            vv[i] = vj - uj * dtx * (v2 - v1 - Qsum * (uj * dt + fp - fm));
        }
    }

    /* reconstruct (cliffs.f:255-262) */
    for (k = 1; k <= kseg; k++) {
        i1 = lghost[k] + 1;
        i2 = rghost[k] - 1;
        for (i = i1; i <= i2; i++) {
            h[i] = (pp[i] - qq[i]) / 4;
            u[i] = (pp[i] + qq[i]) / 2;
        }
    }

    /* shoreline surface-angle limiter (cliffs.f:264-285) */
    if (k1 <= kseg) {
        for (k = k1; k <= kseg; k++) {
            i = lghost[k] + 1;
            j = i + 1;
            if (land[i]) {
                if (SQR2 * h[i] > h[j]) h[i] = h[j] / SQR2;
                if (u[i] < u[j]) u[i] = u[j];
            }
        }
    }
    if (k2 >= 1) {
        for (k = 1; k <= k2; k++) {
            i = rghost[k] - 1;
            j = i - 1;
            if (land[i]) {
                if (SQR2 * h[i] > h[j]) h[i] = h[j] / SQR2;
                if (u[i] > u[j]) u[i] = u[j];
            }
        }
    }

    /* edge integration (cliffs.f:287-329) */
    {
        const int mx = g->mx;
        const double *e1v = g->edge1, *e2v = g->edge2;
        const double eb1_n = e1v[(irw - 1) + (size_t)mx * (iu - 1)];
        const double eb1_t = e1v[(irw - 1) + (size_t)mx * (3 - iu - 1)];
        const double eb1_h = e1v[(irw - 1) + (size_t)mx * 2];
        const double eb2_n = e2v[(irw - 1) + (size_t)mx * (iu - 1)];
        const double eb2_t = e2v[(irw - 1) + (size_t)mx * (3 - iu - 1)];
        const double eb2_h = e2v[(irw - 1) + (size_t)mx * 2];
        double pj, qj, p1, q1, p2, q2, sphr, Qpj, Qjm;

        flood = eb1_h + depth[1];
        if (!newland[1] && !newland[2] && (flood > ground)) {
            pj = Pinv[1];
            qj = Qinv[1];
            p2 = Pinv[2];
            q2 = Qinv[2];
            sphr = (pj - qj) * (pj + qj) * scl[1];
            Qpj = (3 * (q2 + qj) + p2 + pj) * (q2 - qj) / 8 - GRAV * (depth[2] - depth[1]);
            qq[1] = qj - dt * (Qpj / dx[1] + sphr);
            pp[1] = eb1_n + 2 * sqrt(GRAV * flood);
            if (u[1] < 0)
                vv[1] = v[1] - dt * (v[2] - v[1]) * u[1] / dx[1];
            else
                vv[1] = eb1_t;
            h[1] = (pp[1] - qq[1]) / 4;
            u[1] = (pp[1] + qq[1]) / 2;
        }
        flood = depth[n] + eb2_h;
        if (!newland[n - 1] && !newland[n] && (flood > ground)) {
            pj = Pinv[n];
            qj = Qinv[n];
            p1 = Pinv[n - 1];
            q1 = Qinv[n - 1];
            sphr = (pj - qj) * (pj + qj) * scl[n];
            Qjm = (3 * (p1 + pj) + q1 + qj) * (pj - p1) / 8 - GRAV * (depth[n] - depth[n - 1]);
            pp[n] = pj - dt * (Qjm / dx[n - 1] - sphr);
            qq[n] = eb2_n - 2 * sqrt(GRAV * flood);
            if (u[n] > 0)
                vv[n] = v[n] - dt * (v[n] - v[n - 1]) * u[n] / dx[n - 1];
            else
                vv[n] = eb2_t;
            h[n] = (pp[n] - qq[n]) / 4;
            u[n] = (pp[n] + qq[n]) / 2;
        }
    }

    /* scatter (cliffs.f:331-343) */
    if (iu == 1) {
        const size_t row = (size_t)nx * (size_t)(irw - 1);
        for (i = 1; i <= n; i++) {
            g->cel[row + i - 1] = h[i];
            g->xvel[row + i - 1] = u[i];
            g->yvel[row + i - 1] = vv[i];
        }
    } else {
        for (i = 1; i <= n; i++) {
            const size_t o = (size_t)(irw - 1) + (size_t)nx * (i - 1);
            g->cel[o] = h[i];
            g->xvel[o] = vv[i];
            g->yvel[o] = u[i];
        }
    }
    return kseg;
}

int orc_cliffs_line(orc_grid *g, int iu, int irw) {
    orc_ws w;
    int k;
    memset(&w, 0, sizeof(w));
    if (ws_reserve(&w, (iu == 1) ? g->nx : g->ny) != 0) return -2;
    k = cliffs_line_ws(g, iu, irw, &w);
    ws_free(&w);
    return k > 99 ? -1 : k;
}

int orc_sweep(orc_grid *g, int iu, int first, int last) {
    int over = 0;
    const int n = (iu == 1) ? g->nx : g->ny;
#pragma omp parallel reduction(+ : over)
    {
        orc_ws w;
        memset(&w, 0, sizeof(w));
        if (ws_reserve(&w, n) == 0) {
#pragma omp for schedule(static)
            for (int l = first; l <= last; l++) {
                if (cliffs_line_ws(g, iu, l, &w) > 99) over++;
            }
        }
        ws_free(&w);
    }
    return over;
}

void orc_maxwave(orc_grid *g) {
    const size_t m = (size_t)g->nx * (size_t)g->ny;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < m; c++) {
        double aa = g->cel[c];
        if (aa > g->hmax[c]) g->hmax[c] = aa;
        aa = g->xvel[c] * g->xvel[c] + g->yvel[c] * g->yvel[c];
        if (aa > g->velmax[c]) g->velmax[c] = aa;
    }
}

void orc_line_ranges(int nx, int ny, int *ii1, int *ii2, int *jj1, int *jj2,
                     int *xrun, int *yrun, int *mx) {
    int nc, mc;
    if (nx > 2) { *ii1 = 2; *ii2 = nx - 1; *xrun = 1; nc = ny; }
    else { *ii1 = 1; *ii2 = 1; *xrun = 0; nc = 1; }
    if (ny > 2) { *jj1 = 2; *jj2 = ny - 1; *yrun = 1; mc = nx; }
    else { *jj1 = 1; *jj2 = 1; *yrun = 0; mc = 1; }
    *mx = nc > mc ? nc : mc;
}

int orc_steps(orc_grid *g, int nsteps) {
    int ii1, ii2, jj1, jj2, xrun, yrun, mx, over = 0;
    orc_line_ranges(g->nx, g->ny, &ii1, &ii2, &jj1, &jj2, &xrun, &yrun, &mx);
    for (int s = 0; s < nsteps; s++) {
        if (xrun) over += orc_sweep(g, 1, jj1, jj2);
        if (yrun) over += orc_sweep(g, 2, ii1, ii2);
        orc_maxwave(g);
    }
    return over;
}

void orc_depth_to_celerity(int n, const double *H, double ground, double *cel) {
    for (int i = 0; i < n; i++) cel[i] = (H[i] <= ground) ? 0 : sqrt(GRAV * H[i]);
}

void orc_metrics(int nx, int ny, const double *xcrd, const double *ycrd,
                 int cartesian, double *s1, double *s2, double *zeta) {
    const double beta = PI_R * EARTHR / 180.;
    if (nx > 2) {
        for (int j = 0; j < ny; j++) {
            double alpha = cartesian ? 1.0 : cos(PI_R / 180. * ycrd[j]);
            for (int i = 0; i < nx - 1; i++) {
                if (cartesian)
                    s1[(size_t)(nx - 1) * j + i] = xcrd[i + 1] - xcrd[i];
                else
                    s1[(size_t)(nx - 1) * j + i] = alpha * (beta * (xcrd[i + 1] - xcrd[i]));
            }
        }
    }
    if (ny > 2) {
        if (cartesian) {
            for (int i = 0; i < ny - 1; i++) s2[i] = ycrd[i + 1] - ycrd[i];
            for (int i = 0; i < ny; i++) zeta[i] = 0;
        } else {
            const double alpha = 0.125 / EARTHR;
            for (int i = 0; i < ny - 1; i++) s2[i] = beta * (ycrd[i + 1] - ycrd[i]);
            for (int i = 0; i < ny; i++) zeta[i] = alpha * tan(PI_R / 180. * ycrd[i]);
        }
    }
}

void orc_pinpoint(const double *xx, int nx, double xp, int *ip_out, double *cp) {
    int ip = 1, ipp = nx, med;
    while (ip + 2 <= ipp) {
        med = (ip + ipp) / 2;
        if (xp > xx[med - 1]) ip = med;
        else ipp = med;
    }
    if (nx > 1) *cp = (xx[ip] - xp) / (xx[ip] - xx[ip - 1]);
    else *cp = 0.5;
    *ip_out = ip;
}

void orc_nest_maps(int nxp, const double *xp, int nyp, const double *yp,
                   int nxc, const double *xc, int nyc, const double *yc,
                   int *we_ny, double *we_cy, int *sn_nx, double *sn_cx,
                   int *w_nx, double *w_cx, int *e_nx, double *e_cx,
                   int *s_ny, double *s_cy, int *n_ny, double *n_cy) {
    orc_pinpoint(yp, nyp, yc[0], s_ny, s_cy);
    orc_pinpoint(yp, nyp, yc[nyc - 1], n_ny, n_cy);
    for (int i = 0; i < nxc; i++) orc_pinpoint(xp, nxp, xc[i], &sn_nx[i], &sn_cx[i]);
    orc_pinpoint(xp, nxp, xc[0], w_nx, w_cx);
    orc_pinpoint(xp, nxp, xc[nxc - 1], e_nx, e_cx);
    for (int i = 0; i < nyc; i++) orc_pinpoint(yp, nyp, yc[i], &we_ny[i], &we_cy[i]);
}

/* per-node blended value of feed_children_grid.f:56-62 */
static void feed_val(const orc_grid *g, int ii, int kk, double cc, double hmin, double out[3]) {
    const size_t o = (size_t)(ii - 1) + (size_t)g->nx * (kk - 1);
    if (g->dep[o] > hmin) {
        out[0] = cc * g->xvel[o];
        out[1] = cc * g->yvel[o];
        out[2] = cc * (g->cel[o] * g->cel[o] / GRAV - g->dep[o]);
    } else {
        out[0] = out[1] = out[2] = 0;
    }
}

int orc_feed_child(const orc_grid *g, double cuke, double hmin,
                   int nxc, int nyc,
                   const int *we_ny, const double *we_cy,
                   const int *sn_nx, const double *sn_cx,
                   int w_nx, double w_cx, int e_nx, double e_cx,
                   int s_ny, double s_cy, int n_ny, double n_cy,
                   int *iarv,
                   double *uvq_w, double *uvq_e, double *uvq_s, double *uvq_n) {
    const int kk1 = we_ny[0], kk2 = we_ny[nyc - 1] + 1;
    const int ii1 = sn_nx[0], ii2 = sn_nx[nxc - 1] + 1;
    const double thr = (double)0.1f * cuke; /* feed_children_grid.f:30 */
    const int nmax = (g->nx > g->ny ? g->nx : g->ny) + 2;
    double *clm, *clm1;
    if (!*iarv) {
        for (int i = ii1; i <= ii2 && !*iarv; i++) {
            for (int k = kk1; k <= kk2; k++) {
                const size_t o = (size_t)(i - 1) + (size_t)g->nx * (k - 1);
                if (g->dep[o] > hmin) {
                    if (fabs(g->cel[o] * g->cel[o] / GRAV - g->dep[o]) > thr) { *iarv = 1; break; }
                    if ((fabs(g->xvel[o]) + fabs(g->yvel[o])) * g->cel[o] > cuke) { *iarv = 1; break; }
                }
            }
        }
    }
    if (!*iarv) return 0;

    clm = malloc(sizeof(double) * 3 * (size_t)nmax);
    clm1 = malloc(sizeof(double) * 3 * (size_t)nmax);
    if (!clm || !clm1) { free(clm); free(clm1); return -1; }

    /* West and East: blend two parent columns, then along y */
    for (int side = 0; side < 2; side++) {
        const int ii0 = side == 0 ? w_nx : e_nx;
        const double c0 = side == 0 ? w_cx : e_cx;
        double *uvq = side == 0 ? uvq_w : uvq_e;
        for (int kk = kk1; kk <= kk2; kk++) feed_val(g, ii0, kk, c0, hmin, &clm[3 * kk]);
        for (int kk = kk1; kk <= kk2; kk++) feed_val(g, ii0 + 1, kk, 1 - c0, hmin, &clm1[3 * kk]);
        for (int kk = kk1; kk <= kk2; kk++)
            for (int f = 0; f < 3; f++) clm[3 * kk + f] = clm[3 * kk + f] + clm1[3 * kk + f];
        for (int i = 1; i <= nyc; i++) {
            const int kk = we_ny[i - 1];
            const double cc = we_cy[i - 1];
            for (int f = 0; f < 3; f++)
                uvq[(i - 1) + (size_t)nyc * f] = clm[3 * kk + f] * cc + clm[3 * (kk + 1) + f] * (1 - cc);
        }
    }
    /* South and North: blend two parent rows, then along x */
    for (int side = 0; side < 2; side++) {
        const int kk0 = side == 0 ? s_ny : n_ny;
        const double c0 = side == 0 ? s_cy : n_cy;
        double *uvq = side == 0 ? uvq_s : uvq_n;
        for (int ii = ii1; ii <= ii2; ii++) feed_val(g, ii, kk0, c0, hmin, &clm[3 * ii]);
        for (int ii = ii1; ii <= ii2; ii++) feed_val(g, ii, kk0 + 1, 1 - c0, hmin, &clm1[3 * ii]);
        for (int ii = ii1; ii <= ii2; ii++)
            for (int f = 0; f < 3; f++) clm[3 * ii + f] = clm[3 * ii + f] + clm1[3 * ii + f];
        for (int i = 1; i <= nxc; i++) {
            const int ii = sn_nx[i - 1];
            const double cc = sn_cx[i - 1];
            for (int f = 0; f < 3; f++)
                uvq[(i - 1) + (size_t)nxc * f] = clm[3 * ii + f] * cc + clm[3 * (ii + 1) + f] * (1 - cc);
        }
    }
    free(clm);
    free(clm1);
    return 1;
}

void orc_apply_initial_conditions(int nx, int ny, const double *xq, const double *yq,
                                  const double *qq, int kx, int ky, const double *xp,
                                  const double *yp, double *pp) {
    const double zero = (double)0.1e-20f; /* FP32 literal, apply_initial_conditions.f:12 */
    const double x1 = xq[0] - zero, x2 = xq[nx - 1] + zero;
    const double y1 = yq[0] - zero, y2 = yq[ny - 1] + zero;
#define QQ(i, j) qq[(size_t)((i) - 1) + (size_t)nx * ((j) - 1)]
#define PP(i, j) pp[(size_t)((i) - 1) + (size_t)kx * ((j) - 1)]
    int ii, jj;
    double xcc, ycc, q;
    if (nx > 2 && ny > 2) {
        for (int i = 1; i <= kx; i++) {
            const double xx = xp[i - 1];
            if (!(xx >= x1 && xx <= x2)) continue;
            orc_pinpoint(xq, nx, xx, &ii, &xcc);
            for (int j = 1; j <= ky; j++) {
                const double yy = yp[j - 1];
                if (!(yy >= y1 && yy <= y2)) continue;
                orc_pinpoint(yq, ny, yy, &jj, &ycc);
                q = (QQ(ii, jj) * xcc + QQ(ii + 1, jj) * (1 - xcc)) * ycc +
                    (QQ(ii, jj + 1) * xcc + QQ(ii + 1, jj + 1) * (1 - xcc)) * (1 - ycc);
                PP(i, j) = PP(i, j) + q;
            }
        }
    }
    if (ny == 1 && nx > 2) {
        for (int i = 1; i <= kx; i++) {
            const double xx = xp[i - 1];
            if (!(xx >= x1 && xx <= x2)) continue;
            orc_pinpoint(xq, nx, xx, &ii, &xcc);
            q = QQ(ii, 1) * xcc + QQ(ii + 1, 1) * (1 - xcc);
            for (int j = 1; j <= ky; j++) PP(i, j) = PP(i, j) + q;
        }
    }
    if (nx == 1 && ny > 2) {
        for (int j = 1; j <= ky; j++) {
            const double yy = yp[j - 1];
            if (!(yy >= y1 && yy <= y2)) continue;
            orc_pinpoint(yq, ny, yy, &jj, &ycc);
            q = QQ(1, jj) * ycc + QQ(1, jj + 1) * (1 - ycc);
            for (int i = 1; i <= kx; i++) PP(i, j) = PP(i, j) + q;
        }
    }
#undef QQ
#undef PP
}
