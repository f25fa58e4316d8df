// Device building blocks shared by the separate sweep kernels (sweeps.hip)
// and the fused step kernel (step_fused.hip).
//
// x-sweep of one row (cliffs(1,j,nXn), cliffs_main.f:325-329):
//   one wave64 owns a 256-node window of the row (4 contiguous nodes per
//   lane) and produces the middle 248 nodes; the 4-node overlap on each side
//   covers the radius-3 stencil of the wet/dry rules, so every neighbour is
//   one DPP wave shift away (wave_shr/wave_shl) and no LDS is needed.
// y-sweep of one column (cliffs(2,i,nYn), cliffs_main.f:340-344):
//   one lane owns one column and walks its rows with a register sliding
//   window (rows r-1 .. r+3); lanes of a wave are adjacent columns.
//
// The wet/dry segment scan of cliffs.f:65-123 is evaluated in its local
// form (per-node rules of radius <= 3):
//   land_i   = c_i <= celmin
//   RE_i     = land_i & !land_{i-1} & (eta_{i-1} + d_i > ground) & u_{i-1} >= 0
//   start_i  = !land_i & land_{i-1} & !RE_{i-1}        (start_1 = !land_1)
//   LE_i     = start_{i+1} & (eta_{i+1} + d_i > ground) & u_{i+1} <= 0
//   newland  = land & !RE & !LE
// plus the line-global "no wet node -> line untouched" (cliffs.f:113),
// which arrives as a per-line flag.  Where every node in reach is wet and
// deeper than `ground`, the rules are skipped (fast path); the result is
// identical.
#pragma once

#include <hip/hip_runtime.h>

#include "cliffs_dev.h"

namespace cliffs {

// Neighbour lane values by DPP wave shifts; bound_ctrl writes 0 into the lane
// that has no source (lane 0 from the left, lane 63 from the right).
__device__ __forceinline__ float from_left(float x) {  // lane t <- lane t-1
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ float from_right(float x) {  // lane t <- lane t+1
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ int from_left_i(int x) {
    return __builtin_amdgcn_mov_dpp(x, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ int from_right_i(int x) {
    return __builtin_amdgcn_mov_dpp(x, 0x130, 0xF, 0xF, true);
}

// N consecutive nodes held by one lane (N = 4: a 256-node window per wave,
// N = 2: a 128-node window); neighbour shifts along the row.
template <int N>
struct QN {
    float v[N];
};
using Q4 = QN<4>;
template <int N>
__device__ __forceinline__ QN<N> lft(const QN<N>& a) {
    QN<N> r;
    r.v[0] = from_left(a.v[N - 1]);
#pragma unroll
    for (int k = 1; k < N; k++) r.v[k] = a.v[k - 1];
    return r;
}
template <int N>
__device__ __forceinline__ QN<N> rgt(const QN<N>& a) {
    QN<N> r;
#pragma unroll
    for (int k = 0; k + 1 < N; k++) r.v[k] = a.v[k + 1];
    r.v[N - 1] = from_right(a.v[0]);
    return r;
}
// bit k of the lane's nodes shifted by one node along the row
template <int N>
__device__ __forceinline__ int lbits(int b) {
    return ((b << 1) & ((1 << N) - 2)) | ((from_left_i(b) >> (N - 1)) & 1);
}
template <int N>
__device__ __forceinline__ int rbits(int b) {
    return ((b >> 1) & ((1 << (N - 1)) - 1)) | ((from_right_i(b) & 1) << (N - 1));
}
__device__ __forceinline__ bool bit(int m, int k) { return (m >> k) & 1; }

template <int N>
__device__ __forceinline__ QN<N> ldq(const float* p) {
    QN<N> q;
    if constexpr (N == 4) {
        const float4 t = *reinterpret_cast<const float4*>(p);
        q.v[0] = t.x, q.v[1] = t.y, q.v[2] = t.z, q.v[3] = t.w;
    } else {
        const float2 t = *reinterpret_cast<const float2*>(p);
        q.v[0] = t.x, q.v[1] = t.y;
    }
    return q;
}
template <int N>
__device__ __forceinline__ void stq(float* p, const QN<N>& q) {
    if constexpr (N == 4)
        *reinterpret_cast<float4*>(p) = make_float4(q.v[0], q.v[1], q.v[2], q.v[3]);
    else
        *reinterpret_cast<float2*>(p) = make_float2(q.v[0], q.v[1]);
}
template <int N>
__device__ __forceinline__ QN<N> zeroq() {
    QN<N> q;
#pragma unroll
    for (int k = 0; k < N; k++) q.v[k] = 0.f;
    return q;
}
__device__ __forceinline__ Q4 ld4(const float* p) { return ldq<4>(p); }
__device__ __forceinline__ void st4(float* p, const Q4& q) { stq<4>(p, q); }
__device__ __forceinline__ Q4 zero4() { return zeroq<4>(); }

__device__ __forceinline__ bool wave_all(bool x) { return __ballot(!x) == 0; }

// ---------------------------------------------------------------------------
// x-sweep of one row

// Per-lane part of the x window: the lane holds nodes i0 .. i0+N-1; mt points
// at their column metric (XArgs::xm), in registers or in LDS.
struct XLane {
    int i0;
    int vmask;          // bit k: node i0+k lies in 0 .. nx-1
    const float4* mt;   // mt[k]: metric of node i0+k (zero outside the grid)
};

template <int N = 4>
__device__ __forceinline__ int xlane_vmask(int nx, int i0) {
    int m = 0;
#pragma unroll
    for (int k = 0; k < N; k++) m |= (i0 + k >= 0 && i0 + k < nx) ? (1 << k) : 0;
    return m;
}

// General path of x_row: wet/dry front, shorelines, line ends.
template <int N>
__device__ __forceinline__ void x_row_general(const XLane& L, const Params& p, int nx, int ny,
                                               int j, float2 rr, const float* edges, const QN<N>& s,
                                               const QN<N>& cp, const QN<N>& u, const QN<N>& v, const QN<N>& c,
                                               const QN<N>& ge, int land, int deep, QN<N>& ocp, QN<N>& ou,
                                               QN<N>& ov) {
    float ubw = 0.f, vbw = 0.f, gbw = 0.f, ube = 0.f, vbe = 0.f, gbe = 0.f;
    if (edges) {
        ubw = edges[j];
        vbw = edges[ny + j];
        gbw = edges[2 * ny + j];
        ube = edges[3 * ny + j];
        vbe = edges[4 * ny + j];
        gbe = edges[5 * ny + j];
    }
    const QN<N> Lu = lft(u), Lv = lft(v), Lg = lft(ge);
    const int Lland = lbits<N>(land);
    int RE = 0;
#pragma unroll
    for (int k = 0; k < N; k++)
        RE |= (bit(land, k) && !bit(Lland, k) && floods(Lg.v[k], s.v[k], p) && Lu.v[k] >= 0.0f)
                  ? (1 << k)
                  : 0;
    const int start = ~land & Lland & ~lbits<N>(RE) & ((1 << N) - 1);
    const int Rstart = rbits<N>(start);
    const QN<N> Ru = rgt(u), Rv = rgt(v), Rg = rgt(ge);
    int LE = 0;
#pragma unroll
    for (int k = 0; k < N; k++)
        LE |= (bit(L.vmask, k) && bit(Rstart, k) && floods(Rg.v[k], s.v[k], p) && Ru.v[k] <= 0.0f)
                  ? (1 << k)
                  : 0;
    const int nl = land & ~RE & ~LE & ((1 << N) - 1);
    QN<N> pcp, pu, pv, pc, pg;
#pragma unroll
    for (int k = 0; k < N; k++) {
        const float c0 = c0_of(s.v[k]);
        const bool re = bit(RE, k), le = bit(LE, k), n0 = bit(nl, k);
        pcp.v[k] = (re || le) ? p.celmin - c0 : (n0 ? -c0 : cp.v[k]);
        pu.v[k] = re ? Lu.v[k] : (le ? Ru.v[k] : (n0 ? 0.0f : u.v[k]));
        pv.v[k] = re ? Lv.v[k] : (le ? Rv.v[k] : (n0 ? 0.0f : v.v[k]));
        pc.v[k] = c0 + pcp.v[k];
        pg.v[k] = geta_of(s.v[k], pcp.v[k]);
    }
    const int Lnl = lbits<N>(nl), Rnl = rbits<N>(nl);
    const QN<N> Rs = rgt(s), Rpcp = rgt(pcp), Rpu = rgt(pu), Rpv = rgt(pv), Rpc = rgt(pc), Rpg = rgt(pg);
    const QN<N> Ls = lft(s), Lpcp = lft(pcp), Lpu = lft(pu), Lpv = lft(pv), Lpc = lft(pc);
    QN<N> Ep, Eq;
#pragma unroll
    for (int k = 0; k < N; k++) {
        const EdgeRes e = edge_res(Node{s.v[k], pcp.v[k], pu.v[k], pv.v[k]}, pc.v[k], pg.v[k],
                                   Node{Rs.v[k], Rpcp.v[k], Rpu.v[k], Rpv.v[k]}, Rpc.v[k], Rpg.v[k]);
        Ep.v[k] = e.ep;
        Eq.v[k] = e.eq;
    }
    const QN<N> LEp = lft(Ep), LEq = lft(Eq);
    QN<N> lc, lu;
#pragma unroll
    for (int k = 0; k < N; k++) {
        const int i = L.i0 + k;
        const Node X{s.v[k], pcp.v[k], pu.v[k], pv.v[k]};
        const Node Lr{Ls.v[k], Lpcp.v[k], Lpu.v[k], Lpv.v[k]};
        const Node Rr{Rs.v[k], Rpcp.v[k], Rpu.v[k], Rpv.v[k]};
        float xc = pcp.v[k], xu = pu.v[k], xv = pv.v[k];
        const bool edge_node = (i == 0) || (i == nx - 1);
        if (!bit(nl, k) && i >= 1 && i <= nx - 2) {
            const bool gl = bit(Lnl, k), gr = bit(Rnl, k);
            const Node Ln = gl ? mirror(X) : Lr;
            const Node Rn = gr ? mirror(X) : Rr;
            const float cL = gl ? pc.v[k] : Lpc.v[k];
            const float cR = gr ? pc.v[k] : Rpc.v[k];
            const EdgeRes em = gl ? ghost_left_res(pc.v[k], X.u) : EdgeRes{LEp.v[k], LEq.v[k]};
            const EdgeRes ep = gr ? ghost_right_res(pc.v[k], X.u) : EdgeRes{Ep.v[k], Eq.v[k]};
            const bool W = bit(deep, k) && is_deep(Ln.s, p) && is_deep(Rn.s, p);
            Metric m;
            m.rfm = L.mt[k].x * rr.x;
            m.rfp = L.mt[k].y * rr.x;
            m.rsum = L.mt[k].z * rr.x;
            m.fpm = L.mt[k].w * rr.y;
            m.scl = 0.0f;
            interior_update<false>(Ln, cL, X, pc.v[k], Rn, cR, em, ep, W, m, p, xc, xu, xv);
        } else if (!bit(nl, k) && i == 0) {
            if (!bit(Rnl, k) && floods(gbw, X.s, p))
                edge_first(X, pc.v[k], Rr, EdgeRes{Ep.v[k], Eq.v[k]}, L.mt[k].y * rr.x, 0.0f, ubw,
                           vbw, gbw, p, xc, xu, xv);
        } else if (!bit(nl, k) && i == nx - 1) {
            if (!bit(Lnl, k) && floods(gbe, X.s, p))
                edge_last(X, pc.v[k], Lr, EdgeRes{LEp.v[k], LEq.v[k]}, L.mt[k].x * rr.x, 0.0f, ube,
                          vbe, gbe, p, xc, xu, xv);
        }
        ocp.v[k] = xc;
        ou.v[k] = xu;
        ov.v[k] = xv;
        // limiter neighbour source: end nodes contribute pre-step values
        lc.v[k] = c0_of(s.v[k]) + (edge_node ? pcp.v[k] : xc);
        lu.v[k] = edge_node ? pu.v[k] : xu;
    }
    const QN<N> Llc = lft(lc), Llu = lft(lu), Rlc = rgt(lc), Rlu = rgt(lu);
#pragma unroll
    for (int k = 0; k < N; k++) {
        const int i = L.i0 + k;
        if (i >= 1 && i <= nx - 2) {
            if (bit(LE, k) && bit(Lnl, k)) limiter(s.v[k], ocp.v[k], ou.v[k], Rlc.v[k], Rlu.v[k], true, p);
            if (bit(RE, k) && bit(Rnl, k)) limiter(s.v[k], ocp.v[k], ou.v[k], Llc.v[k], Llu.v[k], false, p);
        }
    }
}

// New (cp, u, v) of row j held by the wave as (s, cp, u, v).  `sweep` (wave
// uniform) is false for rows the sweep leaves untouched; they are returned
// unchanged.  rr = (1/a(j), a(j)); edges: [6][ny] west ub, vb, g*eta, east
// ub, vb, g*eta (or null for open edges at rest).
template <int N = 4>
__device__ __forceinline__ void x_row(const XLane& L, const Params& p, int nx, int ny, int j,
                                      bool sweep, float2 rr, const float* edges, const QN<N>& s,
                                      const QN<N>& cp, const QN<N>& u, const QN<N>& v, QN<N>& ocp, QN<N>& ou,
                                      QN<N>& ov) {
    ocp = cp;
    ou = u;
    ov = v;
    if (!sweep) return;
    QN<N> c, ge;
    int land = 0, deep = 0;
#pragma unroll
    for (int k = 0; k < N; k++) {
        c.v[k] = c0_of(s.v[k]) + cp.v[k];
        ge.v[k] = geta_of(s.v[k], cp.v[k]);
        land |= (!bit(L.vmask, k) || is_land(c.v[k], p)) ? (1 << k) : 0;
        deep |= (bit(L.vmask, k) && is_deep(s.v[k], p)) ? (1 << k) : 0;
    }
    if (wave_all(land == 0 && deep == (1 << N) - 1)) {
        // ---- fast path: no land, all deep, no line end in the window
        const QN<N> Rs = rgt(s), Rcp = rgt(cp), Ru = rgt(u), Rv = rgt(v), Rc = rgt(c), Rg = rgt(ge);
        QN<N> Ep, Eq;
#pragma unroll
        for (int k = 0; k < N; k++) {
            const EdgeRes e = edge_res(Node{s.v[k], cp.v[k], u.v[k], v.v[k]}, c.v[k], ge.v[k],
                                       Node{Rs.v[k], Rcp.v[k], Ru.v[k], Rv.v[k]}, Rc.v[k], Rg.v[k]);
            Ep.v[k] = e.ep;
            Eq.v[k] = e.eq;
        }
        const QN<N> LEp = lft(Ep), LEq = lft(Eq);
        const QN<N> Ls = lft(s), Lcp = lft(cp), Lu = lft(u), Lv = lft(v), Lc = lft(c);
#pragma unroll
        for (int k = 0; k < N; k++) {
            Metric m;
            m.rfm = L.mt[k].x * rr.x;
            m.rfp = L.mt[k].y * rr.x;
            m.rsum = L.mt[k].z * rr.x;
            m.fpm = L.mt[k].w * rr.y;
            m.scl = 0.0f;
            interior_update<false>(Node{Ls.v[k], Lcp.v[k], Lu.v[k], Lv.v[k]}, Lc.v[k],
                            Node{s.v[k], cp.v[k], u.v[k], v.v[k]}, c.v[k],
                            Node{Rs.v[k], Rcp.v[k], Ru.v[k], Rv.v[k]}, Rc.v[k],
                            EdgeRes{LEp.v[k], LEq.v[k]}, EdgeRes{Ep.v[k], Eq.v[k]}, true, m, p,
                            ocp.v[k], ou.v[k], ov.v[k]);
        }
        return;
    }
    x_row_general<N>(L, p, nx, ny, j, rr, edges, s, cp, u, v, c, ge, land, deep, ocp, ou, ov);
}

// Bits k of the lane's nodes that are wet in (s, cp) and selected by mask.
template <int N = 4>
__device__ __forceinline__ int wet_bits(const QN<N>& s, const QN<N>& cp, int mask, const Params& p) {
    int wet = 0;
#pragma unroll
    for (int k = 0; k < N; k++)
        wet |= (bit(mask, k) && !is_land(c0_of(s.v[k]) + cp.v[k], p)) ? (1 << k) : 0;
    return wet;
}

// ---------------------------------------------------------------------------
// y-sweep of one column, one row at a time

// Raw state of one row of the lane's column.  In the y-sweep the normal
// velocity u is yvel and the tangential v is xvel (cliffs.f:49-50).
struct RawIn {
    float s, cp, u, v;
};

// Final state of one row of the column, as returned by YPipe::advance.
struct YOut {
    float cp, x, y;  // celerity perturbation, x-velocity, y-velocity
    float s;
};

// Per-column y-sweep constants.
struct YCol {
    bool swept;            // column swept (range and wet flag)
    float e[6];            // south ub, vb, g*eta, north ub, vb, g*eta (0 without edges)
    __device__ __forceinline__ float edge(int k) const { return e[k]; }
    __device__ __forceinline__ void load_edges(const float* edges, int nx, int col, bool ok) {
#pragma unroll
        for (int k = 0; k < 6; k++) e[k] = (edges && ok) ? edges[k * nx + col] : 0.f;
    }
};

// Register pipeline of one column.  The rows are fed in order, one per
// advance(); advance(r, ...) takes row r+3 and returns the final state of
// row r.  Feeding starts at row j0-3 (with r = j0-6) for outputs from row
// j0 on: rows before j0-3 are never needed (stencil radius 3), and the
// values produced for rows < j0 are not final.  All lanes of a wave must
// call advance together (the fast-path test is a wave vote).
//
// The per-row wet/dry decisions carried from row to row are bits of one
// integer (kept in a vector register) rather than separate booleans, which
// the compiler would hold as 64-bit lane masks in scalar registers.
struct YPipe {
    enum : unsigned {
        kLandB = 1u << 0,   // row r+2 is land (or outside the grid)
        kValidB = 1u << 1,  // row r+2 is inside the grid
        kReB = 1u << 2,     // RE of row r+2
        kNl1 = 1u << 3,     // newland of row r+1
        kNl0 = 1u << 4,     //   ... of row r
        kNlm1 = 1u << 5,    //   ... of row r-1
        kLe1 = 1u << 6,     // LE of row r+1
        kRe1 = 1u << 7,     // RE of row r+1
        kLe0 = 1u << 8,     // LE of row r
        kRe0 = 1u << 9,     // RE of row r
    };
    unsigned f;                               // decision bits (above)
    float sB, cpB, uB, vB, cB, geB;           // row r+2 as loaded (+ c, g*eta)
    float uA, vA;                             // row r+1 as loaded (u, v)
    float pcp0, pu0, pv0, pcp1, pu1, pv1;     // loaded state of rows r, r+1
    Node pe1, pe0;                            // post-expansion rows r+1, r
    float pc1, pg1, pc0;
    EdgeRes E0;                               // residual rows r -> r+1
    float ocp0, ou0, ov0, s0;                 // updated row r
    float lc0, lu0, lcm1, lum1;               // limiter sources of rows r, r-1
    int okrun;                                // consecutive wet, deep rows fed

    __device__ __forceinline__ bool has(unsigned b) const { return (f & b) != 0u; }

    __device__ __forceinline__ void init() {
        f = kLandB | kNl1 | kNl0 | kNlm1;
        sB = cpB = uB = vB = cB = geB = 0.f;
        uA = vA = 0.f;
        pcp0 = pu0 = pv0 = pcp1 = pu1 = pv1 = 0.f;
        pe1 = pe0 = Node{0.f, 0.f, 0.f, 0.f};
        pc1 = pg1 = pc0 = 0.f;
        E0 = EdgeRes{0.f, 0.f};
        ocp0 = ou0 = ov0 = s0 = 0.f;
        lc0 = lu0 = lcm1 = lum1 = 0.f;
        okrun = 0;
    }

    // ym/zeta hold the y metric and zeta of rows yb, yb+1, ... (global
    // tables with yb = 0, or a tile's rows staged in LDS)
    __device__ __forceinline__ YOut advance(int r, const RawIn& rin, bool rin_valid, const YCol& col,
                                            int ny, const float4* ym, const float* zeta, int yb,
                                            const Params& p) {
        const bool swept = col.swept;
        // 1. new row r+3
        const float cC = c0_of(rin.s) + rin.cp;
        const float geC = geta_of(rin.s, rin.cp);
        const bool landC = !rin_valid || is_land(cC, p);
        okrun = (!landC && is_deep(rin.s, p)) ? okrun + 1 : 0;
        const int row1 = r + 1;
        // fast path: every swept lane has wet, deep rows r-1..r+3 and row r+1
        // is interior -- no expansion, ghost, limiter or edge can apply
        const bool fast = row1 >= 1 && row1 <= ny - 2 && wave_all(!swept || okrun >= 5);
        unsigned nf;  // decision bits of the next row window
        Node peB;
        float pcB, pgB;
        bool nlB;
        if (fast) {
            nf = 0u;
            nlB = false;
            peB = Node{sB, cpB, uB, vB};
            pcB = cB;
            pgB = geB;
        } else {
            const bool landB = has(kLandB), reB = has(kReB);
            const bool reC = landC && !landB && floods(geB, rin.s, p) && uB >= 0.0f;
            const bool startC = !landC && landB && !reB;
            // 2. expansion decisions and post-expansion state of row r+2
            const bool leB = has(kValidB) && startC && floods(geC, sB, p) && rin.u <= 0.0f;
            nlB = landB && !reB && !leB;
            const float c0B = c0_of(sB);
            peB.s = sB;
            peB.cp = (reB || leB) ? p.celmin - c0B : (nlB ? -c0B : cpB);
            peB.u = reB ? uA : (leB ? rin.u : (nlB ? 0.0f : uB));
            peB.v = reB ? vA : (leB ? rin.v : (nlB ? 0.0f : vB));
            pcB = c0B + peB.cp;
            pgB = geta_of(peB.s, peB.cp);
            nf = (reC ? kReB : 0u) | (nlB ? kNl1 : 0u) | (leB ? kLe1 : 0u) | (reB ? kRe1 : 0u);
        }
        const bool nl1 = has(kNl1), nl0 = has(kNl0);
        // 3. residual across rows r+1 -> r+2, update of row r+1
        const EdgeRes E1 = edge_res(pe1, pc1, pg1, peB, pcB, pgB);
        float ocp1 = pe1.cp, ou1 = pe1.u, ov1 = pe1.v;
        const bool end1 = (row1 == 0) || (row1 == ny - 1);
        if (fast) {
            if (swept) {
                const float4 mr = ym[row1 - yb];
                Metric m;
                m.rfm = mr.x;
                m.rfp = mr.y;
                m.rsum = mr.z;
                m.fpm = mr.w;
                m.scl = zeta[row1 - yb];
                interior_update(pe0, pc0, pe1, pc1, peB, pcB, E0, E1, true, m, p, ocp1, ou1, ov1);
            }
        } else if (swept && !nl1 && row1 >= 0 && row1 < ny) {
            if (row1 >= 1 && row1 <= ny - 2) {
                const float4 mr = ym[row1 - yb];
                Metric m;
                m.rfm = mr.x;
                m.rfp = mr.y;
                m.rsum = mr.z;
                m.fpm = mr.w;
                m.scl = zeta[row1 - yb];
                const Node L = nl0 ? mirror(pe1) : pe0;
                const Node R = nlB ? mirror(pe1) : peB;
                const float cL = nl0 ? pc1 : pc0;
                const float cR = nlB ? pc1 : pcB;
                const EdgeRes em = nl0 ? ghost_left_res(pc1, pe1.u) : E0;
                const EdgeRes ep = nlB ? ghost_right_res(pc1, pe1.u) : E1;
                const bool W = is_deep(pe1.s, p) && is_deep(L.s, p) && is_deep(R.s, p);
                interior_update(L, cL, pe1, pc1, R, cR, em, ep, W, m, p, ocp1, ou1, ov1);
            } else if (row1 == 0) {  // south edge
                const float gbs = col.edge(2);
                if (!nlB && floods(gbs, pe1.s, p))
                    edge_first(pe1, pc1, peB, E1, ym[0 - yb].y, zeta[0 - yb], col.edge(0), col.edge(1),
                               gbs, p, ocp1, ou1, ov1);
            } else {  // north edge
                const float gbn = col.edge(5);
                if (!nl0 && floods(gbn, pe1.s, p))
                    edge_last(pe1, pc1, pe0, E0, ym[ny - 1 - yb].x, zeta[ny - 1 - yb], col.edge(3),
                              col.edge(4), gbn, p, ocp1, ou1, ov1);
            }
        }
        const float lc1 = c0_of(pe1.s) + (end1 ? pe1.cp : ocp1);
        const float lu1 = end1 ? pe1.u : ou1;

        // 4. final state of row r: limiter; unswept columns carry the loaded state
        YOut o;
        if (swept) {
            if (r >= 1 && r <= ny - 2) {
                if (has(kLe0) && has(kNlm1)) limiter(s0, ocp0, ou0, lc1, lu1, true, p);
                if (has(kRe0) && nl1) limiter(s0, ocp0, ou0, lcm1, lum1, false, p);
            }
            o.cp = ocp0;
            o.x = ov0;
            o.y = ou0;
        } else {
            o.cp = pcp0;
            o.x = pv0;
            o.y = pu0;
        }
        o.s = s0;

        // 5. rotate: nl1 -> nl0 -> nlm1, le1/re1 -> le0/re0
        f = nf | ((f << 1) & (kNl0 | kNlm1)) | ((f & (kLe1 | kRe1)) << 2) |
            (landC ? kLandB : 0u) | (rin_valid ? kValidB : 0u);
        lcm1 = lc0;
        lum1 = lu0;
        lc0 = lc1;
        lu0 = lu1;
        ocp0 = ocp1;
        ou0 = ou1;
        ov0 = ov1;
        s0 = pe1.s;
        pe0 = pe1;
        pc0 = pc1;
        pe1 = peB;
        pc1 = pcB;
        pg1 = pgB;
        E0 = E1;
        uA = uB;
        vA = vB;
        pcp0 = pcp1;
        pu0 = pu1;
        pv0 = pv1;
        pcp1 = cpB;
        pu1 = uB;
        pv1 = vB;
        sB = rin.s;
        cpB = rin.cp;
        uB = rin.u;
        vB = rin.v;
        cB = cC;
        geB = geC;
        return o;
    }
};

}  // namespace cliffs
