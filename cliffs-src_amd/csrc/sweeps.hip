// VTCS-2 directional sweeps for CDNA4 (gfx950).
//
// x-sweep (cliffs(1,j,nXn) for all rows, cliffs_main.f:325-329):
//   one wave64 owns a 256-node window of one row (4 contiguous nodes per
//   lane, float4 loads/stores) and writes the middle 248 nodes; the 4-node
//   overlap on each side covers the radius-3 stencil of the wet/dry rules,
//   so every neighbour is one DPP wave shift away (wave_shr/wave_shl) and no
//   LDS is needed.  Each wave walks kXRows rows.
// y-sweep (cliffs(2,i,nYn) for all columns, cliffs_main.f:340-344):
//   one lane owns one column and walks kYRows rows with a register sliding
//   window (rows r-1 .. r+3); lanes of a wave are 64 adjacent columns, so
//   every row access is a coalesced 256-byte segment.  The max-wave update
//   of cliffs_main.f:347-355 is fused into the store.
//
// The wet/dry segment scan of cliffs.f:65-123 is evaluated in its local
// form (per-node rules of radius <= 3):
//   land_i   = c_i <= celmin
//   RE_i     = land_i & !land_{i-1} & (eta_{i-1} + d_i > ground) & u_{i-1} >= 0
//   start_i  = !land_i & land_{i-1} & !RE_{i-1}        (start_1 = !land_1)
//   LE_i     = start_{i+1} & (eta_{i+1} + d_i > ground) & u_{i+1} <= 0
//   newland  = land & !RE & !LE
// plus the line-global "no wet node -> line untouched" (cliffs.f:113),
// which arrives as a per-line flag produced by the previous kernel.
// Where a wave's whole window is wet and deeper than `ground`, the rules
// are skipped (fast path); the result is identical.
#include <hip/hip_runtime.h>

#include "cliffs_dev.h"

namespace cliffs {

__device__ __forceinline__ float from_left(float x) {  // lane t <- lane t-1
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right(float x) {  // lane t <- lane t+1
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ int from_left_i(int x) {
    return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int from_right_i(int x) {
    return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xF, 0xF, false);
}

// Four consecutive nodes held by one lane; neighbour shifts along the row.
struct Q4 {
    float v[4];
};
__device__ __forceinline__ Q4 lft(const Q4& a) {
    return Q4{{from_left(a.v[3]), a.v[0], a.v[1], a.v[2]}};
}
__device__ __forceinline__ Q4 rgt(const Q4& a) {
    return Q4{{a.v[1], a.v[2], a.v[3], from_right(a.v[0])}};
}
__device__ __forceinline__ int lbits(int b) { return ((b << 1) & 0xE) | ((from_left_i(b) >> 3) & 1); }
__device__ __forceinline__ int rbits(int b) { return ((b >> 1) & 0x7) | ((from_right_i(b) & 1) << 3); }
__device__ __forceinline__ bool bit(int m, int k) { return (m >> k) & 1; }

__device__ __forceinline__ Q4 ld4(const float* p) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    return Q4{{t.x, t.y, t.z, t.w}};
}
__device__ __forceinline__ void st4(float* p, const Q4& q) {
    *reinterpret_cast<float4*>(p) = make_float4(q.v[0], q.v[1], q.v[2], q.v[3]);
}
__device__ __forceinline__ Q4 zero4() { return Q4{{0.f, 0.f, 0.f, 0.f}}; }

__device__ __forceinline__ bool wave_all(bool x) { return __ballot(!x) == 0; }

__device__ __forceinline__ void clear_words(unsigned* w, int n, int tid, int nthreads) {
    for (int k = tid; k < n; k += nthreads) w[k] = 0u;
}

// ---------------------------------------------------------------------------
// x-sweep
__global__ __launch_bounds__(256) void sweep_x_kernel(const XArgs a) {
    if (blockIdx.x == 0) clear_words(a.fclr, a.nclr, threadIdx.x, blockDim.x);

    const Params& p = a.p;
    const int lane = threadIdx.x & 63;
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int strip = g % a.nstrips;
    const int chunk = g / a.nstrips;
    const int r0 = a.row_lo + chunk * kXRows;
    if (r0 > a.row_hi) return;
    const int r1 = min(r0 + kXRows - 1, a.row_hi);
    const int nx = a.g.nx, ny = a.g.ny;
    const int i0 = strip * kXStrip - 4 + 4 * lane;  // node held in slot 0
    const bool out_lane = lane >= 1 && lane <= 62;
    const bool ld_lane = i0 >= 0 && i0 < nx;  // lane 0 of strip 0 holds nodes -4..-1: nothing to load

    int vmask = 0;
    float4 mt[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int i = i0 + k;
        const bool ok = i >= 0 && i < nx;
        vmask |= ok ? (1 << k) : 0;
        mt[k] = ok ? a.xm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const int outmask = out_lane ? vmask : 0;
    int colwet = 0;

    // the next row's loads are issued before the current row is computed
    Q4 ns = zero4(), ncp = zero4(), nu = zero4(), nv = zero4();
    auto load_row = [&](int j) {
        if (ld_lane) {
            const size_t ro = (size_t)j * a.g.pitch + i0;
            ns = ld4(a.g.S + ro);
            ncp = ld4(a.g.CP + ro);
            nu = ld4(a.g.U + ro);
            nv = ld4(a.g.V + ro);
        }
    };
    load_row(r0);
    for (int j = r0; j <= r1; ++j) {
        const size_t ro = (size_t)j * a.g.pitch;
        const bool active = j >= a.act_lo && j <= a.act_hi;
        const bool sweep = active && a.fin[j];

        const Q4 s = ns, cp = ncp, u = nu, v = nv;
        if (j < r1) load_row(j + 1);
        Q4 c, ge;
        int land = 0, deep = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            c.v[k] = c0_of(s.v[k]) + cp.v[k];
            ge.v[k] = geta_of(s.v[k], cp.v[k]);
            land |= (!bit(vmask, k) || is_land(c.v[k], p)) ? (1 << k) : 0;
            deep |= (bit(vmask, k) && is_deep(s.v[k], p)) ? (1 << k) : 0;
        }

        Q4 ocp = cp, ou = u, ov = v;
        if (sweep) {
            const float2 rr = a.xr[j];
            if (wave_all(land == 0 && deep == 0xF)) {
                // ---- fast path: no land, all deep, no line end in the window
                const Q4 Rs = rgt(s), Rcp = rgt(cp), Ru = rgt(u), Rv = rgt(v), Rc = rgt(c), Rg = rgt(ge);
                Q4 Ep, Eq;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const EdgeRes e = edge_res(Node{s.v[k], cp.v[k], u.v[k], v.v[k]}, c.v[k], ge.v[k],
                                               Node{Rs.v[k], Rcp.v[k], Ru.v[k], Rv.v[k]}, Rc.v[k],
                                               Rg.v[k]);
                    Ep.v[k] = e.ep;
                    Eq.v[k] = e.eq;
                }
                const Q4 LEp = lft(Ep), LEq = lft(Eq);
                const Q4 Ls = lft(s), Lcp = lft(cp), Lu = lft(u), Lv = lft(v), Lc = lft(c);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    Metric m;
                    m.rfm = mt[k].x * rr.x;
                    m.rfp = mt[k].y * rr.x;
                    m.rsum = mt[k].z * rr.x;
                    m.fpm = mt[k].w * rr.y;
                    m.scl = 0.0f;
                    interior_update(Node{Ls.v[k], Lcp.v[k], Lu.v[k], Lv.v[k]}, Lc.v[k],
                                    Node{s.v[k], cp.v[k], u.v[k], v.v[k]}, c.v[k],
                                    Node{Rs.v[k], Rcp.v[k], Ru.v[k], Rv.v[k]}, Rc.v[k],
                                    EdgeRes{LEp.v[k], LEq.v[k]}, EdgeRes{Ep.v[k], Eq.v[k]}, true, m, p,
                                    ocp.v[k], ou.v[k], ov.v[k]);
                }
            } else {
                // ---- general path: wet/dry front, shorelines, line ends
                float ubw = 0.f, vbw = 0.f, gbw = 0.f, ube = 0.f, vbe = 0.f, gbe = 0.f;
                if (a.edges) {
                    ubw = a.edges[j];
                    vbw = a.edges[ny + j];
                    gbw = a.edges[2 * ny + j];
                    ube = a.edges[3 * ny + j];
                    vbe = a.edges[4 * ny + j];
                    gbe = a.edges[5 * ny + j];
                }
                const Q4 Lu = lft(u), Lv = lft(v), Lg = lft(ge);
                const int Lland = lbits(land);
                int RE = 0;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    RE |= (bit(land, k) && !bit(Lland, k) && floods(Lg.v[k], s.v[k], p) &&
                           Lu.v[k] >= 0.0f)
                              ? (1 << k)
                              : 0;
                const int start = ~land & Lland & ~lbits(RE) & 0xF;
                const int Rstart = rbits(start);
                const Q4 Ru = rgt(u), Rv = rgt(v), Rg = rgt(ge);
                int LE = 0;
#pragma unroll
                for (int k = 0; k < 4; k++)
                    LE |= (bit(vmask, k) && bit(Rstart, k) && floods(Rg.v[k], s.v[k], p) &&
                           Ru.v[k] <= 0.0f)
                              ? (1 << k)
                              : 0;
                const int nl = land & ~RE & ~LE & 0xF;
                Q4 pcp, pu, pv, pc, pg;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float c0 = c0_of(s.v[k]);
                    const bool re = bit(RE, k), le = bit(LE, k), n0 = bit(nl, k);
                    pcp.v[k] = (re || le) ? p.celmin - c0 : (n0 ? -c0 : cp.v[k]);
                    pu.v[k] = re ? Lu.v[k] : (le ? Ru.v[k] : (n0 ? 0.0f : u.v[k]));
                    pv.v[k] = re ? Lv.v[k] : (le ? Rv.v[k] : (n0 ? 0.0f : v.v[k]));
                    pc.v[k] = c0 + pcp.v[k];
                    pg.v[k] = geta_of(s.v[k], pcp.v[k]);
                }
                const int Lnl = lbits(nl), Rnl = rbits(nl);
                const Q4 Rs = rgt(s), Rpcp = rgt(pcp), Rpu = rgt(pu), Rpv = rgt(pv), Rpc = rgt(pc),
                         Rpg = rgt(pg);
                const Q4 Ls = lft(s), Lpcp = lft(pcp), Lpu = lft(pu), Lpv = lft(pv), Lpc = lft(pc);
                Q4 Ep, Eq;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const EdgeRes e = edge_res(Node{s.v[k], pcp.v[k], pu.v[k], pv.v[k]}, pc.v[k],
                                               pg.v[k], Node{Rs.v[k], Rpcp.v[k], Rpu.v[k], Rpv.v[k]},
                                               Rpc.v[k], Rpg.v[k]);
                    Ep.v[k] = e.ep;
                    Eq.v[k] = e.eq;
                }
                const Q4 LEp = lft(Ep), LEq = lft(Eq);
                Q4 lc, lu;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int i = i0 + k;
                    const Node X{s.v[k], pcp.v[k], pu.v[k], pv.v[k]};
                    const Node Lr{Ls.v[k], Lpcp.v[k], Lpu.v[k], Lpv.v[k]};
                    const Node Rr{Rs.v[k], Rpcp.v[k], Rpu.v[k], Rpv.v[k]};
                    float xc = pcp.v[k], xu = pu.v[k], xv = pv.v[k];
                    const bool edge_node = (i == 0) || (i == nx - 1);
                    if (!bit(nl, k) && i >= 1 && i <= nx - 2) {
                        const bool gl = bit(Lnl, k), gr = bit(Rnl, k);
                        const Node L = gl ? mirror(X) : Lr;
                        const Node R = gr ? mirror(X) : Rr;
                        const float cL = gl ? pc.v[k] : Lpc.v[k];
                        const float cR = gr ? pc.v[k] : Rpc.v[k];
                        const EdgeRes em = gl ? ghost_left_res(pc.v[k], X.u) : EdgeRes{LEp.v[k], LEq.v[k]};
                        const EdgeRes ep = gr ? ghost_right_res(pc.v[k], X.u) : EdgeRes{Ep.v[k], Eq.v[k]};
                        const bool W = bit(deep, k) && is_deep(L.s, p) && is_deep(R.s, p);
                        Metric m;
                        m.rfm = mt[k].x * rr.x;
                        m.rfp = mt[k].y * rr.x;
                        m.rsum = mt[k].z * rr.x;
                        m.fpm = mt[k].w * rr.y;
                        m.scl = 0.0f;
                        interior_update(L, cL, X, pc.v[k], R, cR, em, ep, W, m, p, xc, xu, xv);
                    } else if (!bit(nl, k) && i == 0) {
                        if (!bit(Rnl, k) && floods(gbw, X.s, p))
                            edge_first(X, pc.v[k], Rr, EdgeRes{Ep.v[k], Eq.v[k]}, mt[k].y * rr.x, 0.0f,
                                       ubw, vbw, gbw, p, xc, xu, xv);
                    } else if (!bit(nl, k) && i == nx - 1) {
                        if (!bit(Lnl, k) && floods(gbe, X.s, p))
                            edge_last(X, pc.v[k], Lr, EdgeRes{LEp.v[k], LEq.v[k]}, mt[k].x * rr.x, 0.0f,
                                      ube, vbe, gbe, p, xc, xu, xv);
                    }
                    ocp.v[k] = xc;
                    ou.v[k] = xu;
                    ov.v[k] = xv;
                    // limiter neighbour source: end nodes contribute pre-step values
                    lc.v[k] = c0_of(s.v[k]) + (edge_node ? pcp.v[k] : xc);
                    lu.v[k] = edge_node ? pu.v[k] : xu;
                }
                const Q4 Llc = lft(lc), Llu = lft(lu), Rlc = rgt(lc), Rlu = rgt(lu);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int i = i0 + k;
                    if (i >= 1 && i <= nx - 2) {
                        if (bit(LE, k) && bit(Lnl, k))
                            limiter(s.v[k], ocp.v[k], ou.v[k], Rlc.v[k], Rlu.v[k], true, p);
                        if (bit(RE, k) && bit(Rnl, k))
                            limiter(s.v[k], ocp.v[k], ou.v[k], Llc.v[k], Llu.v[k], false, p);
                    }
                }
            }
        }
        // every visited row is written to the other buffer: swept rows with
        // their new values, the others (boundary rows, dry rows) unchanged
        if (out_lane && ld_lane) {
            st4(a.g.CPn + ro + i0, ocp);
            st4(a.g.Un + ro + i0, ou);
            st4(a.g.Vn + ro + i0, ov);
        }

        // wet bits of the new state for the flags consumed by the next sweep
        int wet = 0;
#pragma unroll
        for (int k = 0; k < 4; k++)
            wet |= (bit(outmask, k) && !is_land(c0_of(s.v[k]) + ocp.v[k], p)) ? (1 << k) : 0;
        if (a.flag_rows) {
            if (__ballot(wet != 0) != 0 && lane == 0) a.fout[j] = 1;
        } else {
            colwet |= wet;
        }
        if (a.fuse_max) {
            if (out_lane && ld_lane) {  // fused max-wave (cliffs_main.f:347-355)
                float* HMr = a.g.HM + ro;
                float* VMr = a.g.VM + ro;
                Q4 hm = ld4(HMr + i0), vm = ld4(VMr + i0);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    hm.v[k] = fmaxf(hm.v[k], ocp.v[k]);
                    vm.v[k] = fmaxf(vm.v[k], ou.v[k] * ou.v[k] + ov.v[k] * ov.v[k]);
                }
                st4(HMr + i0, hm);
                st4(VMr + i0, vm);
            }
        }
    }

    if (!a.flag_rows && colwet) {
        if (colwet == 0xF) {
            *reinterpret_cast<unsigned*>(a.fout + i0) = 0x01010101u;
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (bit(colwet, k)) a.fout[i0 + k] = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// y-sweep

// Raw state of one row of the lane's column as loaded from HBM.
struct RawIn {
    float s, cp, u, v;
};

struct Raw {
    float s, cp, u, v, c, ge;
    bool land, valid, deep;
};

__device__ __forceinline__ RawIn load_in(const YArgs& a, int col, int row, bool ok) {
    RawIn r{0.f, 0.f, 0.f, 0.f};
    if (ok) {
        const size_t o = (size_t)row * a.g.pitch + col;
        r.s = a.g.S[o];
        r.cp = a.g.CP[o];
        r.u = a.g.V[o];  // normal velocity of the y-sweep is yvel (cliffs.f:49-50)
        r.v = a.g.U[o];
    }
    return r;
}

__device__ __forceinline__ Raw derive(const RawIn& in, bool valid, const Params& p) {
    Raw r;
    r.valid = valid;
    r.s = in.s;
    r.cp = in.cp;
    r.u = in.u;
    r.v = in.v;
    r.c = c0_of(r.s) + r.cp;
    r.ge = geta_of(r.s, r.cp);
    r.land = !valid || is_land(r.c, p);
    r.deep = valid && is_deep(r.s, p);
    return r;
}

// Rows are processed in batches of kYBatch; the loads of the next batch (state
// three rows ahead, max-wave fields of the rows being finalised) are issued
// before the current batch is computed, so each lane keeps 8 rows of loads in
// flight instead of stalling on one row at a time.
constexpr int kYBatch = 8;

__global__ __launch_bounds__(256) void sweep_y_kernel(const YArgs a) {
    if (blockIdx.x == 0 && blockIdx.y == 0) clear_words(a.fclr, a.nclr, threadIdx.x, blockDim.x);

    const Params& p = a.p;
    const int nx = a.g.nx, ny = a.g.ny;
    const int col = a.col_base + blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool col_ok = col < nx;
    const bool swept = col_ok && col >= a.col_lo && col <= a.col_hi && a.fin[col];
    const int j0 = blockIdx.y * kYRows;
    const int j1 = min(j0 + kYRows, ny);

    float ubs = 0.f, vbs = 0.f, gbs = 0.f, ubn = 0.f, vbn = 0.f, gbn = 0.f;
    if (a.edges && col_ok) {
        ubs = a.edges[col];
        vbs = a.edges[nx + col];
        gbs = a.edges[2 * nx + col];
        ubn = a.edges[3 * nx + col];
        vbn = a.edges[4 * nx + col];
        gbn = a.edges[5 * nx + col];
    }
    auto row_ok = [&](int row) { return col_ok && row >= 0 && row < ny; };

    // pipeline registers (see header): rows relative to iteration r
    Raw rB = derive(load_in(a, col, j0 - 2, row_ok(j0 - 2)), row_ok(j0 - 2), p);  // r+2
    float uA, vA;                                     // r+1 (u, v)
    bool reB;                                         // RE_{r+2}
    float pcp0, pu0, pv0, pcp1, pu1, pv1;             // loaded state of rows r, r+1
    {
        const Raw rA = derive(load_in(a, col, j0 - 3, row_ok(j0 - 3)), row_ok(j0 - 3), p);
        uA = rA.u;
        vA = rA.v;
        reB = rB.land && !rA.land && floods(rA.ge, rB.s, p) && rA.u >= 0.0f;
        pcp0 = pu0 = pv0 = 0.f;
        pcp1 = rA.cp;
        pu1 = rA.u;
        pv1 = rA.v;
    }
    Node pe1{0.f, 0.f, 0.f, 0.f}, pe0{0.f, 0.f, 0.f, 0.f};
    float pc1 = 0.f, pg1 = 0.f, pc0 = 0.f;
    bool nl1 = true, nl0 = true, nlm1 = true;
    bool le1 = false, re1 = false, le0 = false, re0 = false;
    EdgeRes E0{0.f, 0.f};
    float ocp0 = 0.f, ou0 = 0.f, ov0 = 0.f, s0 = 0.f;
    float lc0 = 0.f, lu0 = 0.f, lcm1 = 0.f, lum1 = 0.f;
    unsigned rowwet = 0u;  // lane l: bit0 -> row j0+l, bit1 -> row j0+64+l
    int okrun = 0;         // consecutive loaded rows that are wet and deeper than ground
    bool colwet = false;

    // one iteration: new row r+3 (rin), max-wave fields of row r (hm, vm)
    auto iterate = [&](int r, const RawIn& rin, float hm, float vm) {
        // 1. new row r+3
        const Raw rC = derive(rin, row_ok(r + 3), p);
        okrun = (rC.valid && !rC.land && rC.deep) ? okrun + 1 : 0;
        const int row1 = r + 1;
        // fast path: every swept lane has wet, deep rows r-1..r+3 and row r+1
        // is interior -- no expansion, ghost, limiter or edge can apply
        const bool fast = row1 >= 1 && row1 <= ny - 2 && wave_all(!swept || okrun >= 5);
        bool reC, leB, nlB;
        Node peB;
        float pcB, pgB;
        if (fast) {
            reC = leB = nlB = false;
            peB = Node{rB.s, rB.cp, rB.u, rB.v};
            pcB = rB.c;
            pgB = rB.ge;
        } else {
            reC = rC.land && !rB.land && floods(rB.ge, rC.s, p) && rB.u >= 0.0f;
            const bool startC = !rC.land && rB.land && !reB;
            // 2. expansion decisions and post-expansion state of row r+2
            leB = rB.valid && startC && floods(rC.ge, rB.s, p) && rC.u <= 0.0f;
            nlB = rB.land && !reB && !leB;
            const float c0B = c0_of(rB.s);
            peB.s = rB.s;
            peB.cp = (reB || leB) ? p.celmin - c0B : (nlB ? -c0B : rB.cp);
            peB.u = reB ? uA : (leB ? rC.u : (nlB ? 0.0f : rB.u));
            peB.v = reB ? vA : (leB ? rC.v : (nlB ? 0.0f : rB.v));
            pcB = c0B + peB.cp;
            pgB = geta_of(peB.s, peB.cp);
        }
        // 3. residual across rows r+1 -> r+2, update of row r+1
        const EdgeRes E1 = edge_res(pe1, pc1, pg1, peB, pcB, pgB);
        float ocp1 = pe1.cp, ou1 = pe1.u, ov1 = pe1.v;
        const bool end1 = (row1 == 0) || (row1 == ny - 1);
        if (fast) {
            if (swept) {
                const float4 ym = a.ym[row1];
                Metric m;
                m.rfm = ym.x;
                m.rfp = ym.y;
                m.rsum = ym.z;
                m.fpm = ym.w;
                m.scl = a.zeta[row1];
                interior_update(pe0, pc0, pe1, pc1, peB, pcB, E0, E1, true, m, p, ocp1, ou1, ov1);
            }
        } else if (swept && !nl1 && row1 >= 0 && row1 < ny) {
            if (row1 >= 1 && row1 <= ny - 2) {
                const float4 ym = a.ym[row1];
                Metric m;
                m.rfm = ym.x;
                m.rfp = ym.y;
                m.rsum = ym.z;
                m.fpm = ym.w;
                m.scl = a.zeta[row1];
                const Node L = nl0 ? mirror(pe1) : pe0;
                const Node R = nlB ? mirror(pe1) : peB;
                const float cL = nl0 ? pc1 : pc0;
                const float cR = nlB ? pc1 : pcB;
                const EdgeRes em = nl0 ? ghost_left_res(pc1, pe1.u) : E0;
                const EdgeRes ep = nlB ? ghost_right_res(pc1, pe1.u) : E1;
                const bool W = is_deep(pe1.s, p) && is_deep(L.s, p) && is_deep(R.s, p);
                interior_update(L, cL, pe1, pc1, R, cR, em, ep, W, m, p, ocp1, ou1, ov1);
            } else if (row1 == 0) {
                if (!nlB && floods(gbs, pe1.s, p))
                    edge_first(pe1, pc1, peB, E1, a.ym[0].y, a.zeta[0], ubs, vbs, gbs, p, ocp1, ou1,
                               ov1);
            } else {
                if (!nl0 && floods(gbn, pe1.s, p))
                    edge_last(pe1, pc1, pe0, E0, a.ym[ny - 1].x, a.zeta[ny - 1], ubn, vbn, gbn, p,
                              ocp1, ou1, ov1);
            }
        }
        const float lc1 = c0_of(pe1.s) + (end1 ? pe1.cp : ocp1);
        const float lu1 = end1 ? pe1.u : ou1;

        // 4. finalise row r: limiter, store, max-wave, flags
        if (r >= j0 && r < j1) {
            if (swept && r >= 1 && r <= ny - 2) {
                if (le0 && nlm1) limiter(s0, ocp0, ou0, lc1, lu1, true, p);
                if (re0 && nl1) limiter(s0, ocp0, ou0, lcm1, lum1, false, p);
            }
            const size_t o = (size_t)r * a.g.pitch + col;
            // final cp, x-velocity, y-velocity; boundary and dry columns are
            // carried over unchanged
            const float fcp = swept ? ocp0 : pcp0;
            const float fx = swept ? ov0 : pv0;
            const float fy = swept ? ou0 : pu0;
            bool wet = false;
            if (col_ok) {
                a.g.CPn[o] = fcp;
                a.g.Un[o] = fx;
                a.g.Vn[o] = fy;
                if (a.fuse_max) {  // fused max-wave (cliffs_main.f:347-355): store on change only
                    const float sp2 = fx * fx + fy * fy;
                    if (fcp > hm) a.g.HM[o] = fcp;
                    if (sp2 > vm) a.g.VM[o] = sp2;
                }
                wet = !is_land(c0_of(s0) + fcp, p);
                colwet |= wet;
            }
            const bool anyw = __ballot(wet) != 0;  // uniform control flow
            const int rel = r - j0;
            if (anyw && (rel & 63) == lane) rowwet |= (rel >= 64) ? 2u : 1u;
        }

        // 5. rotate
        lcm1 = lc0;
        lum1 = lu0;
        lc0 = lc1;
        lu0 = lu1;
        nlm1 = nl0;
        ocp0 = ocp1;
        ou0 = ou1;
        ov0 = ov1;
        s0 = pe1.s;
        le0 = le1;
        re0 = re1;
        le1 = leB;
        re1 = reB;
        pe0 = pe1;
        pc0 = pc1;
        nl0 = nl1;
        pe1 = peB;
        pc1 = pcB;
        pg1 = pgB;
        nl1 = nlB;
        E0 = E1;
        uA = rB.u;
        vA = rB.v;
        pcp0 = pcp1;
        pu0 = pu1;
        pv0 = pv1;
        pcp1 = rB.cp;
        pu1 = rB.u;
        pv1 = rB.v;
        rB = rC;
        reB = reC;
    };

    const int r_first = j0 - 4;
    const int n_it = j1 - r_first;
    auto hv_ok = [&](int r) { return a.fuse_max && col_ok && r >= j0 && r < j1; };
    RawIn cur[kYBatch], nxt[kYBatch];
    float hc[kYBatch], vc[kYBatch], hn[kYBatch], vn[kYBatch];
#pragma unroll
    for (int k = 0; k < kYBatch; k++) {
        const int r = r_first + k;
        cur[k] = load_in(a, col, r + 3, k < n_it && row_ok(r + 3));
        const size_t o = (size_t)max(r, 0) * a.g.pitch + col;
        hc[k] = hv_ok(r) ? a.g.HM[o] : 0.f;
        vc[k] = hv_ok(r) ? a.g.VM[o] : 0.f;
    }
    for (int it = 0; it < n_it; it += kYBatch) {
#pragma unroll
        for (int k = 0; k < kYBatch; k++) {  // prefetch the next batch
            const int r = r_first + it + kYBatch + k;
            nxt[k] = load_in(a, col, r + 3, it + kYBatch + k < n_it && row_ok(r + 3));
            const size_t o = (size_t)max(r, 0) * a.g.pitch + col;
            hn[k] = hv_ok(r) ? a.g.HM[o] : 0.f;
            vn[k] = hv_ok(r) ? a.g.VM[o] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < kYBatch; k++)
            if (it + k < n_it) iterate(r_first + it + k, cur[k], hc[k], vc[k]);
#pragma unroll
        for (int k = 0; k < kYBatch; k++) {
            cur[k] = nxt[k];
            hc[k] = hn[k];
            vc[k] = vn[k];
        }
    }

    if (a.flag_cols) {
        if (colwet && col_ok) a.fout[col] = 1;
    } else {
        if ((rowwet & 1u) && j0 + lane < ny) a.fout[j0 + lane] = 1;
        if ((rowwet & 2u) && j0 + 64 + lane < ny) a.fout[j0 + 64 + lane] = 1;
    }
}

// ---------------------------------------------------------------------------
// line flags from the current state: rows (dir 0) or columns (dir 1)
__global__ void line_flags_kernel(DevGrid g, Params p, unsigned char* fout, int dir, unsigned* fclr,
                                  int nclr) {
    if (blockIdx.x == 0 && blockIdx.y == 0) clear_words(fclr, nclr, threadIdx.x, blockDim.x);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    const size_t o = (size_t)j * g.pitch + i;
    const float s = g.S[o];
    if (!is_land(c0_of(s) + g.CP[o], p)) fout[dir == 0 ? j : i] = 1;
}

// Copy lines first..last (rows if dir 0, columns if dir 1) of the state from
// the written buffer (CPn, Un, Vn) back into the current one (CP, U, V):
// used after a sweep over a subset of lines, which leaves the other lines
// of the written buffer stale.
__global__ void copy_lines_kernel(DevGrid g, int dir, int first, int last) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int i, j;
    if (dir == 0) {
        i = t;
        j = first + (int)blockIdx.y;
        if (i >= g.nx) return;
    } else {
        i = first + t;
        j = blockIdx.y;
        if (i > last) return;
    }
    const size_t o = (size_t)j * g.pitch + i;
    g.CP[o] = g.CPn[o];
    g.U[o] = g.Un[o];
    g.V[o] = g.Vn[o];
}

// max-wave update over all cells (cliffs_main.f:347-355)
__global__ void maxwave_kernel(DevGrid g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    const size_t o = (size_t)j * g.pitch + i;
    const float u = g.U[o], v = g.V[o];
    g.HM[o] = fmaxf(g.HM[o], g.CP[o]);
    g.VM[o] = fmaxf(g.VM[o], u * u + v * v);
}

// ---------------------------------------------------------------------------
// launchers
hipError_t launch_sweep_x(const XArgs& a, hipStream_t st) {
    const int rows = a.row_hi - a.row_lo + 1;
    const int chunks = (rows + kXRows - 1) / kXRows;
    const long waves = (long)a.nstrips * chunks;
    const int blocks = (int)((waves + 3) / 4);
    hipLaunchKernelGGL(sweep_x_kernel, dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_sweep_y(const YArgs& a, hipStream_t st) {
    const dim3 grid((a.col_span + 255) / 256, (a.g.ny + kYRows - 1) / kYRows);
    hipLaunchKernelGGL(sweep_y_kernel, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_line_flags(const DevGrid& g, const Params& p, unsigned char* fout, int dir,
                             unsigned* fclr, int nclr, hipStream_t st) {
    const dim3 grid((g.nx + 255) / 256, g.ny);
    hipLaunchKernelGGL(line_flags_kernel, grid, dim3(256), 0, st, g, p, fout, dir, fclr, nclr);
    return hipGetLastError();
}

hipError_t launch_copy_lines(const DevGrid& g, int dir, int first, int last, hipStream_t st) {
    const int n = last - first + 1;
    const dim3 grid = dir == 0 ? dim3((g.nx + 255) / 256, n) : dim3((n + 255) / 256, g.ny);
    hipLaunchKernelGGL(copy_lines_kernel, grid, dim3(256), 0, st, g, dir, first, last);
    return hipGetLastError();
}

hipError_t launch_maxwave(const DevGrid& g, hipStream_t st) {
    const dim3 grid((g.nx + 255) / 256, g.ny);
    hipLaunchKernelGGL(maxwave_kernel, grid, dim3(256), 0, st, g);
    return hipGetLastError();
}

}  // namespace cliffs
