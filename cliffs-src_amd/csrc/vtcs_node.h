// VTCS-2 node arithmetic in FP32 perturbation form, shared by the HIP sweep
// kernels (device) and the host-side emulator used by the test-suite.
//
// Reference algorithm: cliffs.f:17-346 (Titov-Synolakis VTCS-2 with the
// Tolkova land-water boundary).  The reference carries the celerity
// c = sqrt(g H) in FP64.  Stored in FP32, c loses ~3e-4 m of surface
// elevation per rounding in 4 km of water, so the device stores instead
//
//   s  = sign(d) sqrt(g |d|)        background celerity (read-only)
//   cp = c - max(s, 0)              celerity perturbation
//   u, v                            velocities
//
// and every quantity whose background part cancels analytically is
// rewritten so that only perturbation terms are rounded:
//
//   g*eta       = c^2 - s|s| = cp (2 max(s,0) + cp) + min(s,0)^2
//   E_P(a->b)   = 1/8 (l_a + l_b)(P_b - P_a) - g (d_b - d_a)
//               = A + B,  E_Q(a->b) = A - B, with
//   A = 1/2 (u_a + u_b)(u_b - u_a) + (g eta_b - g eta_a)
//   B = (u_a + u_b)(c_b - c_a) + 1/2 (c_a + c_b)(u_b - u_a)
//
// (l = 3P + Q for the P family, 3Q + P for the Q family; P = u + 2c,
// Q = u - 2c; cliffs.f:133-134, 223-247).  At rest u = 0, cp = 0 and both
// residuals are exactly zero, so the scheme is well balanced in FP32.
// The bathymetry the device integrates over is d' = s|s|/g, which differs
// from the FP64 depth by < 1.2e-7 relative.
//
// Ghost nodes at shorelines (cliffs.f:137-152) are mirrors: (-Q, -P, v, d)
// of the adjacent wet node is the node with u -> -u and c, d unchanged.
//
// Rounding is fixed by the source: every file is compiled with FP
// contraction off, and each fused multiply-add is written as fmaf.  The
// device kernels (fused or not) and the host emulator therefore round every
// operation identically: their FP32 results agree bit for bit.  Divisions
// and square roots are the correctly rounded IEEE operations on both.
#pragma once

#if defined(__HIPCC__)
#define CLIFFS_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define CLIFFS_HD static inline
#endif

namespace cliffs {

// Per-launch physical constants.
struct Params {
    float dt;       // time step
    float grav;     // g = (float)9.8 (global_modules.f:17 widens the same FP32 literal)
    float celmin;   // sqrt(g ground), rounded once from FP64 (cliffs_main.f:227)
    float gground;  // g * ground
    float gn2;      // g * crough  (almost-Manning friction, cliffs.f:204)
    float sqrt2;    // (float)1.41421356 (cliffs.f:34)
    int friction;   // crough > 0
};

// One node as seen by a sweep: u is the velocity normal to the line
// direction of travel (x-velocity in the x-sweep), v the tangential one.
struct Node {
    float s, cp, u, v;
};

// Per-node metric of the line (cliffs.f:215-217, 249-250).
struct Metric {
    float rfm;   // 1 / dx(i-1)
    float rfp;   // 1 / dx(i)
    float rsum;  // 1 / (dx(i-1) + dx(i))
    float fpm;   // dx(i) - dx(i-1)
    float scl;   // zeta(i) in the y-sweep, 0 in the x-sweep
};

CLIFFS_HD float c0_of(float s) { return fmaxf(s, 0.0f); }
CLIFFS_HD float cel_of(const Node& n) { return c0_of(n.s) + n.cp; }
CLIFFS_HD float geta_of(float s, float cp) {
    const float sn = fminf(s, 0.0f);
    return fmaf(cp, fmaf(2.0f, c0_of(s), cp), sn * sn);
}
// g*eta over deep water (s > 0): the background term min(s,0)^2 is 0
CLIFFS_HD float geta_deep(float s, float cp) { return cp * fmaf(2.0f, s, cp); }
// u^2 + v^2 as every kernel and the emulator round it
CLIFFS_HD float speed2(float u, float v) { return fmaf(u, u, v * v); }
CLIFFS_HD bool is_land(float c, const Params& p) { return c <= p.celmin; }
// d > ground  <=>  s > celmin for the device bathymetry (cliffs.f:221)
CLIFFS_HD bool is_deep(float s, const Params& p) { return s > p.celmin; }
// eta_a + d_b > ground, in g-scaled form (cliffs.f:77, 92, 288, 309)
CLIFFS_HD bool floods(float geta_a, float s_b, const Params& p) {
    return fmaf(s_b, fabsf(s_b), geta_a) > p.gground;
}
CLIFFS_HD Node mirror(const Node& n) { return Node{n.s, n.cp, -n.u, n.v}; }

// Characteristic residuals E_P, E_Q across the edge a -> b (see header).
struct EdgeRes {
    float ep, eq;
};
CLIFFS_HD EdgeRes edge_res(const Node& a, float ca, float gea, const Node& b, float cb,
                           float geb) {
    const float U = a.u + b.u;
    const float Cs = ca + cb;
    const float du = b.u - a.u;
    const float dc = (c0_of(b.s) - c0_of(a.s)) + (b.cp - a.cp);
    const float A = fmaf(0.5f * U, du, geb - gea);
    const float B = fmaf(U, dc, 0.5f * Cs * du);
    return EdgeRes{A + B, A - B};
}
// Residuals against the node's own reflective ghost, left and right.
CLIFFS_HD EdgeRes ghost_left_res(float c, float u) {
    const float t = 2.0f * c * u;
    return EdgeRes{t, -t};
}
CLIFFS_HD EdgeRes ghost_right_res(float c, float u) {
    const float t = 2.0f * c * u;
    return EdgeRes{-t, t};
}

// Interior VTCS-2 update of node X between neighbours L and R (real nodes or
// ghosts), cliffs.f:200-260.  em/ep are the residuals L->X and X->R.
// W is (d > ground) for all three nodes.  kScl = false drops the metric term
// (P-Q)(P+Q) zeta, which is zero in the x-sweep (scl = 0, cliffs.f:39).
template <bool kScl = true>
CLIFFS_HD void interior_update(const Node& L, float cL, const Node& X, float cX, const Node& R,
                               float cR, EdgeRes em, EdgeRes ep, bool W, const Metric& m,
                               const Params& p, float& cp_out, float& u_out, float& v_out) {
    const float u = X.u, v = X.v, c = cX;
    const float dtx = p.dt * m.rsum;

    float F = 0.0f;  // almost-Manning friction g n^2 u sqrt((u^2+v^2)/H)/H, H = c^2/g
    if (p.friction) {
        // = g n^2 * g * u * sqrt(g (u^2+v^2)) / c^3   (c >= celmin > 0 here)
        F = (p.gn2 * p.grav * u * sqrtf(p.grav * speed2(u, v))) / (c * c * c);
    }
    const float S = kScl ? 8.0f * c * u * m.scl : 0.0f;  // (P-Q)(P+Q) zeta

    // P family
    const float qpjP = ep.ep * m.rfp;
    const float qjmP = em.ep * m.rfm;
    float qsP = 0.5f * (qpjP + qjmP);
    // Q family
    const float qpjQ = ep.eq * m.rfp;
    const float qjmQ = em.eq * m.rfm;
    float qsQ = 0.5f * (qpjQ + qjmQ);

    // transcritical branch (cliffs.f:228-229, 241-242); rare.  P needs
    // u + c < 0 on the left (e1 = 4(u1+c1) < 0), Q needs u - c > 0 on the
    // right (e2 = 4(u2-c2) > 0): one wave vote on that screen skips the branch
    // in the common case, without a divergent branch.
    const bool screen = (L.u + cL < 0.0f) || (R.u - cR > 0.0f);
#if defined(__HIP_DEVICE_COMPILE__)
    if (__ballot(screen) != 0) {
#else
    if (screen) {
#endif
        const bool tcP = W && (L.u + cL < 0.0f) && (R.u + cR > 0.0f);
        const bool tcQ = W && (L.u - cL < 0.0f) && (R.u - cR > 0.0f);
        const float duRL = R.u - L.u;
        const float dcpRL = R.cp - L.cp;
        const float dsRL = R.s - L.s;
        const float wP = u + c, wQ = u - c;
        const float vP = fmaf(2.0f * (u + X.cp), dsRL, fmaf(2.0f * wP, dcpRL, wP * duRL)) * m.rsum;
        const float vQ = fmaf(-2.0f * (u - X.cp), dsRL, fmaf(-2.0f * wQ, dcpRL, wQ * duRL)) * m.rsum;
        qsP = tcP ? vP : qsP;
        qsQ = tcQ ? vQ : qsQ;
    }
    qsP = fmaf((u + c) * dtx, qjmP - qpjP, qsP);
    qsQ = fmaf((u - c) * dtx, qjmQ - qpjQ, qsQ);

    // P' = P - dt (qsP + F - S), Q' = Q - dt (qsQ + F + S);
    // c' = (P'-Q')/4, u' = (P'+Q')/2
    cp_out = fmaf(-0.25f * p.dt, fmaf(-2.0f, S, qsP - qsQ), X.cp);
    u_out = fmaf(-0.5f * p.dt, fmaf(2.0f, F, qsP + qsQ), u);

    // tangential velocity advection (cliffs.f:249-250)
    const float B = fmaf(R.v - v, m.rfp, -((v - L.v) * m.rfm));
    v_out = fmaf(-(u * dtx), fmaf(-B, fmaf(u, p.dt, m.fpm), R.v - L.v), v);
}

// Incoming characteristic 2 (sqrt(g (eta_b + d)) - max(s,0)) from boundary
// elevation eta_b (cliffs.f:299, 320), without cancellation in deep water.
CLIFFS_HD float incoming_2dc(float geta_b, float s) {
    const float X = fmaf(s, fabsf(s), geta_b);
    const float r = sqrtf(fmaxf(X, 0.0f));
    return 2.0f * ((s > 0.0f) ? geta_b / (r + s) : r);
}

// Open/forced edge at the first node (cliffs.f:288-308).  X = node 1, N =
// node 2 (post-expansion), res = E(1->2).  ub/vb/geb: boundary normal and
// tangential velocity and g*eta.  rdx = 1/dx(1).
CLIFFS_HD void edge_first(const Node& X, float cX, const Node& N, EdgeRes res, float rdx,
                          float scl, float ub, float vb, float geb, const Params& p,
                          float& cp_out, float& u_out, float& v_out) {
    const float S = 8.0f * cX * X.u * scl;
    const float qn = fmaf(-p.dt, fmaf(res.eq, rdx, S), fmaf(-2.0f, X.cp, X.u));
    const float pn = ub + incoming_2dc(geb, X.s);
    cp_out = 0.25f * (pn - qn);
    u_out = 0.5f * (pn + qn);
    v_out = (X.u < 0.0f) ? fmaf(-(p.dt * (N.v - X.v) * X.u), rdx, X.v) : vb;
}

// Open/forced edge at the last node (cliffs.f:309-329).  X = node n, N =
// node n-1, res = E(n-1->n), rdx = 1/dx(n-1).
CLIFFS_HD void edge_last(const Node& X, float cX, const Node& N, EdgeRes res, float rdx,
                         float scl, float ub, float vb, float geb, const Params& p,
                         float& cp_out, float& u_out, float& v_out) {
    const float S = 8.0f * cX * X.u * scl;
    const float pn = fmaf(-p.dt, fmaf(res.ep, rdx, -S), fmaf(2.0f, X.cp, X.u));
    const float qn = ub - incoming_2dc(geb, X.s);
    cp_out = 0.25f * (pn - qn);
    u_out = 0.5f * (pn + qn);
    v_out = (X.u > 0.0f) ? fmaf(-(p.dt * (X.v - N.v) * X.u), rdx, X.v) : vb;
}

// Shoreline surface-angle limiter (cliffs.f:264-285).  For a newly flooded
// end node X with neighbour celerity cN and velocity uN: cap c at cN/sqrt2
// (tested as sqrt2*c > cN, as the reference does) and clamp u.
CLIFFS_HD void limiter(float s, float& cp, float& u, float cN, float uN, bool left,
                       const Params& p) {
    const float c = c0_of(s) + cp;
    if (p.sqrt2 * c > cN) cp = cN / p.sqrt2 - c0_of(s);
    if (left ? (u < uN) : (u > uN)) u = uN;
}

}  // namespace cliffs
