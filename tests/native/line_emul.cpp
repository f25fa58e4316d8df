// Test-only host emulator of the device sweep algorithm.
//
// Runs the same FP32 node arithmetic as the HIP kernels (vtcs_node.h), the
// same per-node metric tables (separable x spacing, FP32 reciprocals built
// as the library builds them) and the same local-rule form of the wet/dry
// front logic, line by line on the CPU.  With contraction off on both sides
// and every fused multiply-add explicit, its FP32 state equals the device's
// bit for bit.  The test-suite uses it to (a) check the local rules against the
// sequential segment scan of cliffs.f:65-123 as the oracle executes it and
// (b) measure the FP32 error budget against the FP64 oracle without a GPU.
// It is not part of the product library and nothing in the package loads it.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../cliffs-src_amd/csrc/convert.h"
#include "../../cliffs-src_amd/csrc/vtcs_node.h"

using namespace cliffs;

namespace {

struct F4 {
    float x, y, z, w;
};

struct Emul {
    int nx, ny;
    std::vector<double> dep;
    std::vector<float> S, CP, U, V, HM, VM;
    // the library's metric tables: per column (1/b(i-1), 1/b(i), 1/(b(i-1)+b(i)),
    // b(i)-b(i-1)), per row (1/a(j), a(j)) with s1(i,j) = a(j) b(i); per row
    // of the y-sweep the same four from s2, and zeta
    std::vector<F4> xm, ym;
    std::vector<float> xra, xrb, zt;
    Params p;
    Consts64 k;
};

struct BC {
    float ub, vb, geb;
};

// One line: nodes nd[0..n-1] (u normal, v tangential), per-node metric
// mt[0..n-1], edge spacings 1/dx(1) and 1/dx(n-1), metric term at the ends.
// Writes results back into nd.  Returns false if the line had no wet node and
// was left untouched (cliffs.f:113).
// kScl: the y-sweep's metric term (the x-sweep drops it, as the kernels do).
template <bool kScl>
bool sweep_line(const Params& p, int n, Node* nd, const Metric* mt, float rdx0, float rdx1,
                float scl0, float scl1, BC bl, BC br) {
    std::vector<float> c(n), ge(n);
    std::vector<char> land(n), RE(n, 0), LE(n, 0), start(n, 0), nl(n);
    bool any = false;
    for (int i = 0; i < n; i++) {
        c[i] = cel_of(nd[i]);
        ge[i] = geta_of(nd[i].s, nd[i].cp);
        land[i] = is_land(c[i], p);
        any |= !land[i];
    }
    if (!any) return false;
    // right expansion RE_i and segment starts
    for (int i = 1; i < n; i++)
        RE[i] = land[i] && !land[i - 1] && floods(ge[i - 1], nd[i].s, p) && nd[i - 1].u >= 0.0f;
    start[0] = !land[0];
    for (int i = 1; i < n; i++) start[i] = !land[i] && land[i - 1] && !RE[i - 1];
    // left expansion LE_i (node i flooded from a segment starting at i+1)
    for (int i = 0; i + 1 < n; i++)
        LE[i] = start[i + 1] && floods(ge[i + 1], nd[i].s, p) && nd[i + 1].u <= 0.0f;
    // post-expansion state
    std::vector<Node> pe(nd, nd + n);
    for (int i = 0; i < n; i++) {
        nl[i] = land[i] && !RE[i] && !LE[i];
        const float c0 = c0_of(nd[i].s);
        if (RE[i]) pe[i] = Node{nd[i].s, p.celmin - c0, nd[i - 1].u, nd[i - 1].v};
        else if (LE[i]) pe[i] = Node{nd[i].s, p.celmin - c0, nd[i + 1].u, nd[i + 1].v};
        else if (nl[i]) pe[i] = Node{nd[i].s, -c0, 0.0f, 0.0f};
    }
    std::vector<float> pc(n), pg(n);
    for (int i = 0; i < n; i++) {
        pc[i] = cel_of(pe[i]);
        pg[i] = geta_of(pe[i].s, pe[i].cp);
    }
    // shared real-edge residuals between i and i+1
    std::vector<EdgeRes> er(n > 1 ? n - 1 : 1);
    for (int i = 0; i + 1 < n; i++) er[i] = edge_res(pe[i], pc[i], pg[i], pe[i + 1], pc[i + 1], pg[i + 1]);

    std::vector<Node> out(pe);
    for (int i = 1; i + 1 < n; i++) {
        if (nl[i]) continue;
        const Node X = pe[i];
        const bool gl = nl[i - 1], gr = nl[i + 1];
        const Node L = gl ? mirror(X) : pe[i - 1];
        const Node R = gr ? mirror(X) : pe[i + 1];
        const float cL = gl ? pc[i] : pc[i - 1], cR = gr ? pc[i] : pc[i + 1];
        const EdgeRes em = gl ? ghost_left_res(pc[i], X.u) : er[i - 1];
        const EdgeRes ep = gr ? ghost_right_res(pc[i], X.u) : er[i];
        const bool W = is_deep(X.s, p) && is_deep(L.s, p) && is_deep(R.s, p);
        interior_update<kScl>(L, cL, X, pc[i], R, cR, em, ep, W, mt[i], p, out[i].cp, out[i].u,
                              out[i].v);
    }
    // limiter: left cases (LE_i, newland_{i-1}), then right (RE_i, newland_{i+1})
    for (int i = 1; i < n; i++) {
        if (LE[i] && nl[i - 1] && i + 1 < n) {
            const Node& J = (i + 1 == n - 1) ? pe[i + 1] : out[i + 1];
            limiter(out[i].s, out[i].cp, out[i].u, cel_of(J), J.u, true, p);
        }
    }
    for (int i = 0; i + 1 < n; i++) {
        if (RE[i] && nl[i + 1] && i >= 1) {
            const Node& J = (i - 1 == 0) ? pe[i - 1] : out[i - 1];
            limiter(out[i].s, out[i].cp, out[i].u, cel_of(J), J.u, false, p);
        }
    }
    // edges
    if (!nl[0] && !nl[1] && floods(bl.geb, pe[0].s, p))
        edge_first(pe[0], pc[0], pe[1], er[0], rdx0, scl0, bl.ub, bl.vb, bl.geb, p, out[0].cp,
                   out[0].u, out[0].v);
    if (!nl[n - 2] && !nl[n - 1] && floods(br.geb, pe[n - 1].s, p))
        edge_last(pe[n - 1], pc[n - 1], pe[n - 2], er[n - 2], rdx1, scl1, br.ub, br.vb, br.geb, p,
                  out[n - 1].cp, out[n - 1].u, out[n - 1].v);
    for (int i = 0; i < n; i++) nd[i] = out[i];
    return true;
}

}  // namespace

extern "C" {

void* emul_create(int nx, int ny, const double* dep, const double* s1, const double* s2,
                  const double* zeta, double dt, double ground, double crough) {
    Emul* e = new Emul;
    e->nx = nx;
    e->ny = ny;
    const size_t m = (size_t)nx * ny;
    e->dep.assign(dep, dep + m);
    e->k = make_consts(ground);
    e->S.resize(m);
    for (size_t i = 0; i < m; i++) e->S[i] = to_s(dep[i], e->k);
    e->CP.assign(m, 0.0f);
    e->U.assign(m, 0.0f);
    e->V.assign(m, 0.0f);
    e->HM.resize(m);
    for (size_t i = 0; i < m; i++) e->HM[i] = -c0f(e->S[i]);
    e->VM.assign(m, 0.0f);
    // metric tables exactly as cliffs_create builds them
    e->xm.assign(nx, F4{0.f, 0.f, 0.f, 0.f});
    e->xra.assign(ny, 1.0f);
    e->xrb.assign(ny, 1.0f);
    if (nx > 2) {
        const int l = nx - 1;
        std::vector<double> b(s1, s1 + l);
        for (int i = 0; i < nx; i++) {
            const double bm = i >= 1 ? b[i - 1] : b[0];
            const double bp = i < l ? b[i] : b[l - 1];
            e->xm[i] = F4{(float)(1.0 / bm), (float)(1.0 / bp), (float)(1.0 / (bm + bp)), (float)(bp - bm)};
        }
        for (int j = 0; j < ny; j++) {
            const double aj = s1[(size_t)l * j] / b[0];
            e->xra[j] = (float)(1.0 / aj);
            e->xrb[j] = (float)aj;
        }
    }
    e->ym.assign(ny, F4{0.f, 0.f, 0.f, 0.f});
    e->zt.assign(ny, 0.0f);
    if (ny > 2) {
        for (int j = 0; j < ny; j++) {
            const double bm = j >= 1 ? s2[j - 1] : s2[0];
            const double bp = j < ny - 1 ? s2[j] : s2[ny - 2];
            e->ym[j] = F4{(float)(1.0 / bm), (float)(1.0 / bp), (float)(1.0 / (bm + bp)), (float)(bp - bm)};
            e->zt[j] = (float)zeta[j];
        }
    }
    e->p.dt = (float)dt;
    e->p.grav = 9.8f;
    e->p.celmin = e->k.celmin32;
    e->p.gground = (float)(e->k.grav * ground);
    e->p.gn2 = (float)(e->k.grav * crough);
    e->p.sqrt2 = 1.41421356f;
    e->p.friction = crough > 0;
    return e;
}

void emul_destroy(void* h) { delete (Emul*)h; }

void emul_upload(void* h, const double* cel, const double* u, const double* v) {
    Emul* e = (Emul*)h;
    const size_t m = (size_t)e->nx * e->ny;
    for (size_t i = 0; i < m; i++) {
        e->CP[i] = to_cp(cel[i], e->dep[i], e->S[i], e->k);
        e->U[i] = (float)u[i];
        e->V[i] = (float)v[i];
    }
}

void emul_download(void* h, double* cel, double* u, double* v, double* hmax, double* velmax) {
    Emul* e = (Emul*)h;
    const size_t m = (size_t)e->nx * e->ny;
    for (size_t i = 0; i < m; i++) {
        cel[i] = to_cel(e->S[i], e->CP[i], e->dep[i], e->k);
        u[i] = e->U[i];
        v[i] = e->V[i];
        hmax[i] = to_cel(e->S[i], e->HM[i], e->dep[i], e->k);
        velmax[i] = e->VM[i];
    }
}

// One sweep of direction iu over lines first..last (1-based) with edge
// arrays (mx,3) column-major FP64 (NULL = zero), as cliffs_main.f:325-344.
void emul_sweep(void* h, int iu, int first, int last, int mx, const double* e1, const double* e2) {
    Emul* e = (Emul*)h;
    const int nx = e->nx, ny = e->ny;
    auto bc = [&](const double* a, int irw) {
        BC b{0, 0, 0};
        if (a) {
            b.ub = (float)a[(irw - 1) + (size_t)mx * (iu - 1)];
            b.vb = (float)a[(irw - 1) + (size_t)mx * (2 - iu)];
            b.geb = (float)(e->k.grav * a[(irw - 1) + (size_t)mx * 2]);
        }
        return b;
    };
    std::vector<Node> nd(nx > ny ? nx : ny);
    std::vector<Metric> mt(nx > ny ? nx : ny);
    if (iu == 1) {
        for (int j = first; j <= last; j++) {
            const size_t row = (size_t)nx * (j - 1);
            for (int i = 0; i < nx; i++) {
                nd[i] = Node{e->S[row + i], e->CP[row + i], e->U[row + i], e->V[row + i]};
                const float ra = e->xra[j - 1];
                mt[i] = Metric{e->xm[i].x * ra, e->xm[i].y * ra, e->xm[i].z * ra,
                               e->xm[i].w * e->xrb[j - 1], 0.0f};
            }
            const float rdx0 = e->xm[0].y * e->xra[j - 1];
            const float rdx1 = e->xm[nx - 1].x * e->xra[j - 1];
            if (sweep_line<false>(e->p, nx, nd.data(), mt.data(), rdx0, rdx1, 0.0f, 0.0f, bc(e1, j),
                           bc(e2, j))) {
                for (int i = 0; i < nx; i++) {
                    e->CP[row + i] = nd[i].cp;
                    e->U[row + i] = nd[i].u;
                    e->V[row + i] = nd[i].v;
                }
            }
        }
    } else {
        for (int i = first; i <= last; i++) {
            for (int j = 0; j < ny; j++) {
                const size_t o = (size_t)(i - 1) + (size_t)nx * j;
                nd[j] = Node{e->S[o], e->CP[o], e->V[o], e->U[o]};
                mt[j] = Metric{e->ym[j].x, e->ym[j].y, e->ym[j].z, e->ym[j].w, e->zt[j]};
            }
            if (sweep_line<true>(e->p, ny, nd.data(), mt.data(), e->ym[0].y, e->ym[ny - 1].x, e->zt[0],
                           e->zt[ny - 1], bc(e1, i), bc(e2, i))) {
                for (int j = 0; j < ny; j++) {
                    const size_t o = (size_t)(i - 1) + (size_t)nx * j;
                    e->CP[o] = nd[j].cp;
                    e->V[o] = nd[j].u;
                    e->U[o] = nd[j].v;
                }
            }
        }
    }
}

void emul_maxwave(void* h) {
    Emul* e = (Emul*)h;
    const size_t m = (size_t)e->nx * e->ny;
    for (size_t c = 0; c < m; c++) {
        if (e->CP[c] > e->HM[c]) e->HM[c] = e->CP[c];
        const float vv = speed2(e->U[c], e->V[c]);
        if (vv > e->VM[c]) e->VM[c] = vv;
    }
}

// nsteps full steps (cliffs_main.f:317-355) with fixed edges.
void emul_steps(void* h, int nsteps, int mx, const double* west, const double* east,
                const double* south, const double* north) {
    Emul* e = (Emul*)h;
    const int nx = e->nx, ny = e->ny;
    const bool xrun = nx > 2, yrun = ny > 2;
    for (int st = 0; st < nsteps; st++) {
        if (xrun) emul_sweep(h, 1, yrun ? 2 : 1, yrun ? ny - 1 : 1, mx, west, east);
        if (yrun) emul_sweep(h, 2, xrun ? 2 : 1, xrun ? nx - 1 : 1, mx, south, north);
        emul_maxwave(h);
    }
}

}  // extern "C"
