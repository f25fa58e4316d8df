// FP64 line mode (include/cliffs_fp64.h): the reference's line procedure
// (cliffs.f:17-346) in double precision on the device, one wave per line.
//
// HBM layout.  The state (c, xvel, yvel), the maxima and the depth are
// (nx, ny) arrays of doubles with x fastest, the reference's own order; a
// y-sweep works on column-major copies, so every line is contiguous.  Each
// wave of a sweep owns one line.  The line's per-node flags and segment
// bounds are contiguous in one int buffer, so the wave's lanes, one node
// each, touch consecutive addresses; the sweep updates the state in place.
//
// Arithmetic.  The reference's expressions in the reference's order, in
// double precision, compiled with FP contraction off (no fused
// multiply-add is formed) and with IEEE division and square root: every
// operation rounds as the reference's FP64 build rounds it.
//
// Parallelism.  Every pass of a line runs on its wave's 64 lanes: the
// segment scan (cliffs.f:65-111), sequential in the reference, as local
// rules per node and a prefix count of the segment starts (see the kernel),
// the integration per node, the limiter per segment.  Lines are
// independent, so a sweep runs them all at once.  It is a fidelity mode for
// flow that FP32 state cannot follow: its FP64 traffic makes it several
// times slower than the FP32 fused step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "../../include/cliffs_fp64.h"
#include "cliffs_dev.h"

namespace {

constexpr double kG = (double)9.8f;            // g (global_modules.f:17, an FP32 literal)
constexpr double kSqr2 = (double)1.41421356f;  // cliffs.f:34

// Per-line work arrays: the per-node flags and the segments' bounds (ints,
// n + 2 slots each), and twelve doubles per line that keep the state of the
// line's first two and last two nodes after the run-up expansions, which
// the edge integration reads once the sweep has overwritten them.
enum : int { kLG, kRG, kFL, kInts };  // segment bounds (by segment); per-node flag bits below
enum : int {
    kLand = 1, kNewland = 2,                         // as loaded; still dry after the expansions
    kR = 4, kW = 8, kS = 16, kL = 32, kM = 64, kE = 128,  // decisions of the segment scan
    kEnd = 256,                                      // a segment ends here (not merged away)
    kIn = 512,                                       // inside a segment: lghost < i < rghost
    kFromR = 1024,                                   // flooded by the start at i+1 (its left expansion)
};
constexpr int kStash = 16;  // doubles per line: (h, u, v) of nodes 1, 2, n-1, n

struct Line64 {
    int nx, iu, first, nlines, n;
    int ld;  // node i of line l at ld*(l-1) + (i-1): nx for rows, ny for the transposed columns
    const double* dep;
    double* cel;
    double* xvel;
    double* yvel;
    const double* s1;
    const double* s2;
    const double* zeta;
    const double* e1;  // edge arrays (mx, 3) column-major, or null (zeros)
    const double* e2;
    int mx;
    double dt, ground, crough, celmin;
    double* wd;  // nlines * kStash
    int* wi;     // nlines * kInts * (n + 2)
};

// Int work arrays of one line: array k, element i at (line * kInts + k) * np + i
// (np = n + 2), so the lanes of the line's wave touch consecutive addresses.
struct Work {
    int* k;
    size_t np;
    __device__ int& ix(int a, int i) const { return k[(size_t)a * np + i]; }
};

// The lanes of a wave read what other lanes of it wrote in the pass before
// (all of one work-group: one L1).
__device__ __forceinline__ void pass_done() { __threadfence_block(); }

// x of lane l, in every lane
__device__ __forceinline__ double lane_value(double x, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, l);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

struct Node { double h, u, v; };  // celerity, normal and tangential velocity

constexpr int kLineWaves = 4;  // lines per block
constexpr int kChunks = 4;     // 64-node chunks whose loads a flag pass keeps in flight

// One wave per line, one node (or segment) per lane in every pass.  Two
// passes over the line's flags decide the segment scan; one pass over the
// state then integrates the nodes in place, 64 at a time, each from its own
// and its neighbours' state after the run-up expansions; the limiter and the
// edges close the sweep.
__global__ __launch_bounds__(64 * kLineWaves) void line64_kernel(const Line64 a) {
    const int lane = threadIdx.x & 63;
    // line of this wave (made scalar: the line's bases then sit in scalar
    // registers and every access is a 32-bit offset from them)
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * kLineWaves + (threadIdx.x >> 6));
    if (t >= a.nlines) return;                                   // wave-uniform
    const int irw = a.first + t;                                 // 1-based line
    const int n = a.n, nx = a.nx, iu = a.iu;
    const double G = kG, celmin = a.celmin, ground = a.ground, dt = a.dt;
    const size_t np = (size_t)n + 2;
    const Work w{a.wi + (size_t)t * kInts * np, np};
    double* const sd = a.wd + (size_t)t * kStash;
    // the line (cliffs.f:36-54): c, normal velocity (xvel in x, yvel in y),
    // tangential velocity, depth, spacing dx(i) between nodes i and i+1, and
    // the metric term (zeta in y, 0 in x)
    double* const cel = a.cel;
    double* const nv = iu == 1 ? a.xvel : a.yvel;
    double* const tv = iu == 1 ? a.yvel : a.xvel;
    const double* const dep = a.dep;
    auto at = [&](int i) { return (size_t)a.ld * (irw - 1) + (i - 1); };
    auto dx = [&](int i) { return iu == 1 ? a.s1[(size_t)(nx - 1) * (irw - 1) + (i - 1)] : a.s2[i - 1]; };
    auto scl = [&](int i) { return iu == 1 ? 0.0 : a.zeta[i - 1]; };

    // Segment scan with run-up expansion (cliffs.f:58-111).  The reference
    // walks the line once; every decision of that walk depends only on a
    // node and its neighbours' original values, because a node that floods
    // (run-up expansion) is land and its wet neighbour is never changed:
    //   R(i)  right expansion: land i, wet i-1, flood from i-1, u(i-1) >= 0
    //   W(i)  the walk is in water after i: !land(i) || R(i)
    //   S(i)  a segment starts: wet i after land i-1 not expanded (i = 1: wet)
    //   L(i)  left expansion at a start: flood from i, u(i) <= 0
    //   M(i)  the start merges with the segment before (cliffs.f:103-107):
    //         S, L and W(i-2), i.e. a single land node between the two
    //   E(i)  the segment ends at land node i: W(i-1) and not R(i)
    // so the lanes decide them node by node, and a prefix count of the
    // starts numbers the segments.  One pass, kChunks chunks of 64 nodes
    // at a time (their loads in flight together): each lane decides land, R
    // and W of its node, takes those of its neighbours from the neighbour
    // lanes (the chunk before's last two nodes, the chunk after's first,
    // carried across chunks), then S, L, M, E of its node and of node i+1
    // (whose left expansion floods i and whose merge removes i's end), the
    // nodes still dry, and the segment numbers: starts not merged, counted
    // along the line; a node is inside a segment (lghost < i < rghost) when
    // more segments have started than ended up to it, or when the start
    // after it floods it, never at the line's ends.
    int kseg = 0, kend = 0;
    double hlast = 0.0;  // c of the node before the batch
    int fl62 = 0, fl63 = 0;  // land, R, W of the two nodes before the batch
    for (int c0 = 0; c0 < n; c0 += 64 * kChunks) {
        double h[kChunks];
        int fa[kChunks], fm1[kChunks], fm2[kChunks], fp1[kChunks];
#pragma unroll
        for (int u = 0; u < kChunks; u++) {
            const int i = c0 + 1 + lane + 64 * u;
            h[u] = i <= n ? cel[at(i)] : 0.0;
        }
        const int inext = c0 + 64 * kChunks + 1;  // the node after the batch
        const bool land_next = inext <= n && cel[at(inext)] <= celmin;
#pragma unroll
        for (int u = 0; u < kChunks; u++) {  // land, R, W
            const int i = c0 + 1 + lane + 64 * u;
            const double left = u == 0 ? hlast : lane_value(h[u - 1], 63);
            const double up = __shfl_up(h[u], 1);
            const double hp = lane == 0 ? left : up;  // c(i-1)
            const bool land = h[u] <= celmin;
            bool r = false;
            if (i >= 2 && i <= n && land && !(hp <= celmin)) {
                const double flood = (hp * hp) / G - dep[at(i - 1)] + dep[at(i)];
                r = (flood > ground) && (nv[at(i - 1)] >= 0);
            }
            fa[u] = (land ? kLand : 0) | (r ? kR : 0) | ((!land || r) ? kW : 0);
        }
#pragma unroll
        for (int u = 0; u < kChunks; u++) {  // the neighbours' land, R, W
            const int i = c0 + 1 + lane + 64 * u;
            const int p62 = u == 0 ? fl62 : __builtin_amdgcn_readlane(fa[u - 1], 62);
            const int p63 = u == 0 ? fl63 : __builtin_amdgcn_readlane(fa[u - 1], 63);
            const int nx0 = u + 1 < kChunks ? __builtin_amdgcn_readlane(fa[u + 1 < kChunks ? u + 1 : u], 0)
                                            : (land_next ? kLand : 0);
            const int up1 = __shfl_up(fa[u], 1), up2 = __shfl_up(fa[u], 2), dn1 = __shfl_down(fa[u], 1);
            fm1[u] = i >= 2 ? (lane == 0 ? p63 : up1) : 0;
            fm2[u] = i >= 3 ? (lane == 0 ? p62 : lane == 1 ? p63 : up2) : 0;
            fp1[u] = i + 1 <= n ? (lane == 63 ? nx0 : dn1) : 0;
        }
        hlast = lane_value(h[kChunks - 1], 63);
        fl62 = __builtin_amdgcn_readlane(fa[kChunks - 1], 62);
        fl63 = __builtin_amdgcn_readlane(fa[kChunks - 1], 63);
#pragma unroll
        for (int u = 0; u < kChunks; u++) {
            const int c = c0 + 64 * u;
            if (c >= n) break;  // wave-uniform
            const int i = c + 1 + lane;
            int fl = 0;
            bool st = false, en = false, fbn = false;
            if (i <= n) {
                fl = fa[u];
                const bool land = fl & kLand;
                const bool s0 = i == 1 ? !land : (!land && (fm1[u] & kLand) && !(fm1[u] & kR));
                bool l0 = false;
                if (s0 && i >= 2) {
                    const double hi = h[u];
                    const double flood = (hi * hi) / G - dep[at(i)] + dep[at(i - 1)];
                    l0 = (flood > ground) && (nv[at(i)] <= 0);
                }
                const bool m0 = s0 && l0 && i >= 3 && (fm2[u] & kW);
                const bool e0 = i >= 2 && land && (fm1[u] & kW) && !(fl & kR);
                const bool sn = i + 1 <= n && !(fp1[u] & kLand) && land && !(fl & kR);
                bool ln = false;
                if (sn) {
                    const double hn = cel[at(i + 1)];
                    const double flood = (hn * hn) / G - dep[at(i + 1)] + dep[at(i)];
                    ln = (flood > ground) && (nv[at(i + 1)] <= 0);
                }
                const bool mn = sn && ln && i >= 2 && (fm1[u] & kW);
                const bool end = e0 && !mn;
                const bool newland = land && !(fl & kR) && !ln;
                fl |= (newland ? kNewland : 0) | (s0 ? kS : 0) | (l0 ? kL : 0) | (m0 ? kM : 0) | (e0 ? kE : 0) |
                      (end ? kEnd : 0) | (ln ? kFromR : 0);
                st = s0 && !m0;
                en = end;
                fbn = sn && !mn && ln;
            }
            const unsigned long long ms = __ballot(st), me = __ballot(en);
            const unsigned long long upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);  // lanes <= this one
            const int below = __popcll(ms & ((1ull << lane) - 1ull));
            if (st) w.ix(kLG, kseg + below + 1) = i == 1 ? 1 : ((fl & kL) ? (i - 2 > 1 ? i - 2 : 1) : i - 1);
            if (en) w.ix(kRG, kseg + below + (st ? 1 : 0)) = i;
            const bool inside = kseg + __popcll(ms & upto) > kend + __popcll(me & upto) || fbn;
            if (i <= n) w.ix(kFL, i) = fl | ((i >= 2 && i <= n - 1 && inside) ? kIn : 0);
            kseg += __popcll(ms);
            kend += __popcll(me);
        }
    }
    if (kseg == 0) return;  // cliffs.f:113: the line is left untouched
    if (kend < kseg && lane == 0) w.ix(kRG, kseg) = n;  // the last segment runs to the line's end
    pass_done();
    const int k1 = (w.ix(kFL, 1) & kLand) ? 1 : 2;
    const int rlast = w.ix(kRG, kseg);
    const int k2 = (w.ix(kFL, rlast) & kNewland) ? kseg : kseg - 1;

    // Node i after the run-up expansions (cliffs.f:81-107: a flooded node
    // takes c_min and the velocities of the wet node that floods it) and the
    // emptying of nodes still dry (cliffs.f:114-124).  The sources are
    // original values: a node that floods is wet and never changed itself.
    auto post = [&](int i) -> Node {
        if (i > n) return {0.0, 0.0, 0.0};
        const int fl = w.ix(kFL, i);
        if (fl & kNewland) return {0.0, 0.0, 0.0};
        if (fl & kFromR) return {celmin, nv[at(i + 1)], tv[at(i + 1)]};
        if (fl & kR) return {celmin, nv[at(i - 1)], tv[at(i - 1)]};
        return {cel[at(i)], nv[at(i)], tv[at(i)]};
    };
    // Flags, depths and spacings of node i and its neighbours.
    struct Aux {
        int fl, flm, flp;
        double dj, d1, d2, fm, fp;
    };
    auto aux = [&](int i) -> Aux {
        Aux x{0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (i <= n) {
            x.fl = w.ix(kFL, i);
            x.dj = dep[at(i)];
            if (i >= 2) x.flm = w.ix(kFL, i - 1), x.fm = dx(i - 1);
            if (i + 1 <= n) x.flp = w.ix(kFL, i + 1), x.fp = dx(i);
            // the neighbours' depths come from the neighbour lanes, but at
            // the chunk's edges
            if (lane == 0 && i >= 2) x.d1 = dep[at(i - 1)];
            if (lane == 63 && i + 1 <= n) x.d2 = dep[at(i + 1)];
        }
        return x;
    };
    // Riemann invariants P = u + 2c, Q = u - 2c inside the segments and at
    // the line's wet end nodes (cliffs.f:125-136, 153-164).  At a segment's
    // first (last) interior node the node before (after) it is a ghost: the
    // line's end node 1 (n) itself when wet (cliffs.f:153-164), else the
    // reflection (-Q, -P, v, d) of the interior node (cliffs.f:137-152).
    // VTCS-2 integration of the interior nodes (cliffs.f:168-253) and c, u
    // from the new invariants (cliffs.f:255-262), 64 nodes at a time, in
    // place: a chunk's own nodes and the next two chunks' nodes are read
    // before the chunk is written, and the previous chunk's last node is
    // kept in registers.
    // The loads run ahead of the integration: the state of the chunk after
    // next, the flags, depths and spacings of the next chunk.
    Node cur = post(1 + lane);
    Node nxt = 64 < n ? post(65 + lane) : Node{0.0, 0.0, 0.0};
    Aux ac = aux(1 + lane);
    Node prev{0.0, 0.0, 0.0};
    for (int c = 0; c < n; c += 64) {
        const int i = c + 1 + lane;
        const Node nx2 = c + 128 < n ? post(i + 128) : Node{0.0, 0.0, 0.0};
        const Aux an = c + 64 < n ? aux(i + 64) : Aux{0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0.0};
        Node lft, rgt;
        lft.h = __shfl_up(cur.h, 1);
        lft.u = __shfl_up(cur.u, 1);
        lft.v = __shfl_up(cur.v, 1);
        rgt.h = __shfl_down(cur.h, 1);
        rgt.u = __shfl_down(cur.u, 1);
        rgt.v = __shfl_down(cur.v, 1);
        const double nh = lane_value(nxt.h, 0), nu = lane_value(nxt.u, 0), nvv = lane_value(nxt.v, 0);
        if (lane == 0) lft = prev;
        if (lane == 63) rgt = Node{nh, nu, nvv};
        const double dl = __shfl_up(ac.dj, 1), dr = __shfl_down(ac.dj, 1);
        if (lane != 0) ac.d1 = dl;
        if (lane != 63) ac.d2 = dr;
        if (i <= n) {
            const int fl = ac.fl;
            double H = cur.h, U = cur.u, VV = cur.v;
            if (fl & kIn) {  // 2 <= i <= n-1
                const double hc = cur.h, uj = cur.u, vj = cur.v;
                const double pj = uj + 2 * hc, qj = uj - 2 * hc;
                const double dj = ac.dj;
                const int flm = ac.flm, flp = ac.flp;
                double p1, q1, v1, d1, p2, q2, v2, d2;
                if ((flm & kIn) || (i - 1 == 1 && !(flm & kNewland))) {
                    p1 = lft.u + 2 * lft.h, q1 = lft.u - 2 * lft.h, v1 = lft.v, d1 = ac.d1;
                } else {
                    p1 = -qj, q1 = -pj, v1 = vj, d1 = dj;
                }
                if ((flp & kIn) || (i + 1 == n && !(flp & kNewland))) {
                    p2 = rgt.u + 2 * rgt.h, q2 = rgt.u - 2 * rgt.h, v2 = rgt.v, d2 = ac.d2;
                } else {
                    p2 = -qj, q2 = -pj, v2 = vj, d2 = dj;
                }
                const double hj = hc * hc / G;
                // almost-Manning friction (cliffs.f:204)
                const double cc = G * a.crough * uj * sqrt((uj * uj + vj * vj) / hj) / hj;
                const double sphr = (pj - qj) * (pj + qj) * scl(i);
                const double fm = ac.fm, fp = ac.fp;
                const double dtx = dt / (fm + fp);
                const double dpj = G * (d2 - dj), dpm = G * (dj - d1);
                const bool wet3 = (dj > ground) && (d1 > ground) && (d2 > ground);

                // P family
                double ej = 3 * pj + qj, e2 = 3 * p2 + q2, e1 = 3 * p1 + q1;
                double Qpj = (0.125 * (e2 + ej) * (p2 - pj) - dpj) / fp;
                double Qjm = (0.125 * (ej + e1) * (pj - p1) - dpm) / fm;
                double Qsum;
                if ((e1 < 0) && (e2 > 0) && wet3)
                    Qsum = (0.25 * ej * (p2 - p1) - 2 * G * (sqrt(dj * d2) - sqrt(dj * d1))) / (fm + fp);
                else
                    Qsum = (Qpj + Qjm) / 2;
                Qsum = Qsum + 0.25 * ej * dtx * (Qjm - Qpj);
                const double pp = pj - dt * (Qsum + cc - sphr);

                // Q family
                ej = 3 * qj + pj, e2 = 3 * q2 + p2, e1 = 3 * q1 + p1;
                Qpj = (0.125 * (e2 + ej) * (q2 - qj) - dpj) / fp;
                Qjm = (0.125 * (ej + e1) * (qj - q1) - dpm) / fm;
                if ((e1 < 0) && (e2 > 0) && wet3)
                    Qsum = (0.25 * ej * (q2 - q1) - 2 * G * (sqrt(dj * d2) - sqrt(dj * d1))) / (fm + fp);
                else
                    Qsum = (Qpj + Qjm) / 2;
                Qsum = Qsum + 0.25 * ej * dtx * (Qjm - Qpj);
                const double qq = qj - dt * (Qsum + cc + sphr);

                // tangential velocity (cliffs.f:249-250)
                Qsum = (v2 - vj) / fp - (vj - v1) / fm;
                VV = vj - uj * dtx * (v2 - v1 - Qsum * (uj * dt + fp - fm));
                H = (pp - qq) / 4;
                U = (pp + qq) / 2;
            }
            // the edge integration reads nodes 1, 2, n-1, n as they were
            if (i <= 2) {
                sd[3 * (i - 1)] = cur.h, sd[3 * (i - 1) + 1] = cur.u, sd[3 * (i - 1) + 2] = cur.v;
            }
            if (i >= n - 1) {
                const int s = i == n ? 9 : 6;
                sd[s] = cur.h, sd[s + 1] = cur.u, sd[s + 2] = cur.v;
            }
            const size_t o = at(i);
            cel[o] = H;
            nv[o] = U;
            tv[o] = VV;
        }
        prev = Node{lane_value(cur.h, 63), lane_value(cur.u, 63), lane_value(cur.v, 63)};
        cur = nxt;
        nxt = nx2;
        ac = an;
    }
    pass_done();
    // shoreline surface-angle limiter on newly flooded end nodes
    // (cliffs.f:264-285): all left ends, then all right ends, one segment
    // per lane (segments are disjoint, so the ends of one pass do not meet)
    for (int k = k1 + lane; k <= kseg; k += 64) {
        const int i = w.ix(kLG, k) + 1, j = i + 1;
        if ((w.ix(kFL, i) & kLand) && j <= n) {
            const size_t oi = at(i), oj = at(j);
            if (kSqr2 * cel[oi] > cel[oj]) cel[oi] = cel[oj] / kSqr2;
            if (nv[oi] < nv[oj]) nv[oi] = nv[oj];
        }
    }
    pass_done();
    for (int k = 1 + lane; k <= k2; k += 64) {
        const int i = w.ix(kRG, k) - 1, j = i - 1;
        if (j >= 1 && (w.ix(kFL, i) & kLand)) {
            const size_t oi = at(i), oj = at(j);
            if (kSqr2 * cel[oi] > cel[oj]) cel[oi] = cel[oj] / kSqr2;
            if (nv[oi] > nv[oj]) nv[oi] = nv[oj];
        }
    }
    pass_done();

    // open or forced edges at the line's ends (cliffs.f:287-329), lane 0,
    // from the invariants of nodes 1, 2 (n-1, n) before the integration
    if (lane == 0) {
        const int mx = a.mx;
        auto edge = [&](const double* e, int col) {
            return e ? e[(size_t)(irw - 1) + (size_t)mx * col] : 0.0;
        };
        const double b1n = edge(a.e1, iu - 1), b1t = edge(a.e1, 3 - iu - 1), b1h = edge(a.e1, 2);
        const double b2n = edge(a.e2, iu - 1), b2t = edge(a.e2, 3 - iu - 1), b2h = edge(a.e2, 2);
        double flood = b1h + dep[at(1)];
        if (!(w.ix(kFL, 1) & kNewland) && !(w.ix(kFL, 2) & kNewland) && (flood > ground)) {
            const double h1 = sd[0], u1 = sd[1], v1 = sd[2], h2 = sd[3], u2 = sd[4], v2 = sd[5];
            const bool in2 = (w.ix(kFL, 2) & kIn) != 0;
            const double pj = u1 + 2 * h1, qj = u1 - 2 * h1;
            const double p2 = in2 ? u2 + 2 * h2 : 0.0, q2 = in2 ? u2 - 2 * h2 : 0.0;
            const double sphr = (pj - qj) * (pj + qj) * scl(1);
            const double Qpj = (3 * (q2 + qj) + p2 + pj) * (q2 - qj) / 8 - G * (dep[at(2)] - dep[at(1)]);
            const double qq = qj - dt * (Qpj / dx(1) + sphr);
            const double pp = b1n + 2 * sqrt(G * flood);
            const size_t o = at(1);
            tv[o] = u1 < 0 ? v1 - dt * (v2 - v1) * u1 / dx(1) : b1t;
            cel[o] = (pp - qq) / 4;
            nv[o] = (pp + qq) / 2;
        }
        flood = dep[at(n)] + b2h;
        if (!(w.ix(kFL, n - 1) & kNewland) && !(w.ix(kFL, n) & kNewland) && (flood > ground)) {
            const double h1 = sd[6], u1 = sd[7], v1 = sd[8], hn = sd[9], un = sd[10], vn = sd[11];
            const bool in1 = (w.ix(kFL, n - 1) & kIn) != 0;
            const double pj = un + 2 * hn, qj = un - 2 * hn;
            const double p1 = in1 ? u1 + 2 * h1 : 0.0, q1 = in1 ? u1 - 2 * h1 : 0.0;
            const double sphr = (pj - qj) * (pj + qj) * scl(n);
            const double Qjm = (3 * (p1 + pj) + q1 + qj) * (pj - p1) / 8 - G * (dep[at(n)] - dep[at(n - 1)]);
            const double pp = pj - dt * (Qjm / dx(n - 1) - sphr);
            const double qq = b2n - 2 * sqrt(G * flood);
            const size_t o = at(n);
            tv[o] = un > 0 ? vn - dt * (vn - v1) * un / dx(n - 1) : b2t;
            cel[o] = (pp - qq) / 4;
            nv[o] = (pp + qq) / 2;
        }
    }
}

// out (cols x rows) = in (rows x cols) transposed, through 32 x 32 LDS tiles
// (the y-sweep runs on the column-major copy, so its lanes read along
// contiguous lines as the x-sweep's do)
__global__ __launch_bounds__(256) void transpose64_kernel(const double* __restrict__ in, double* __restrict__ out,
                                                          int rows, int cols) {
    __shared__ double tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = in[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[tx][k];
    }
}

// The y-sweep's column-major c, xvel, yvel back into the (nx, ny) state,
// through 32 x 32 LDS tiles, and the step's running maxima
// (cliffs_main.f:347-355) from the same values: the sweep and the maxima
// of one step in one pass over the state.
__global__ __launch_bounds__(256) void untranspose_max64_kernel(const double* __restrict__ cT,
                                                                const double* __restrict__ xT,
                                                                const double* __restrict__ yT,
                                                                double* __restrict__ cel, double* __restrict__ xv,
                                                                double* __restrict__ yv, double* __restrict__ hm,
                                                                double* __restrict__ vm, int rows, int cols) {
    __shared__ double tc[32][33], tu[32][33], tw[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) {
            const size_t g = (size_t)r * cols + c;
            tc[k][tx] = cT[g], tu[k][tx] = xT[g], tw[k][tx] = yT[g];
        }
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) {
            const size_t o = (size_t)c * rows + r;
            const double h = tc[tx][k], u = tu[tx][k], v = tw[tx][k];
            cel[o] = h, xv[o] = u, yv[o] = v;
            if (h > hm[o]) hm[o] = h;
            const double aa = u * u + v * v;
            if (aa > vm[o]) vm[o] = aa;
        }
    }
}

// running maxima (cliffs_main.f:347-355)
__global__ void maxwave64_kernel(const double* __restrict__ cel, const double* __restrict__ xv,
                                 const double* __restrict__ yv, double* __restrict__ hm,
                                 double* __restrict__ vm, size_t m) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    double aa = cel[c];
    if (aa > hm[c]) hm[c] = aa;
    aa = xv[c] * xv[c] + yv[c] * yv[c];
    if (aa > vm[c]) vm[c] = aa;
}

// Gauge sampling (output_gages.f:18-29), one thread per gauge.
__global__ void gauges64_kernel(const double* __restrict__ cel, const double* __restrict__ xv,
                                const double* __restrict__ yv, const double* __restrict__ dep, int nx, int ny,
                                double celmin, const int* __restrict__ ij, int n, double* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = ij[q] - 1, j = ij[n + q] - 1;
    const size_t o = (size_t)j * nx + i;
    const double c = cel[o];
    out[q] = c > celmin ? c * c / kG - dep[o] : -1e32;
    out[n + q] = xv[o];
    out[2 * n + q] = yv[o];
}

// Largest Courant number (|u| + c) dt / dx over wet cells (dx the smaller
// spacing next to the cell, along the directions swept) and the count of
// non-finite cells; Courant numbers are >= 0, so their bit patterns order
// as the values do.
__global__ void health64_kernel(const double* __restrict__ cel, const double* __restrict__ xv,
                                const double* __restrict__ yv, const double* __restrict__ s1,
                                const double* __restrict__ s2, int nx, int ny, int xrun, int yrun, double dt,
                                double celmin, unsigned long long* out) {
    const size_t m = (size_t)nx * ny;
    const size_t c0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double cmax = 0.0;
    unsigned long long nbad = 0;
    for (size_t cidx = c0; cidx < m; cidx += (size_t)gridDim.x * blockDim.x) {
        const double c = cel[cidx], u = xv[cidx], v = yv[cidx];
        if (!(isfinite(c) && isfinite(u) && isfinite(v))) {
            nbad++;
            continue;
        }
        if (c <= celmin) continue;
        const int i = (int)(cidx % nx), j = (int)(cidx / nx);
        if (xrun) {
            double rdx = 0.0;
            if (i >= 1) rdx = fmax(rdx, 1.0 / s1[(size_t)(nx - 1) * j + i - 1]);
            if (i <= nx - 2) rdx = fmax(rdx, 1.0 / s1[(size_t)(nx - 1) * j + i]);
            cmax = fmax(cmax, (fabs(u) + c) * dt * rdx);
        }
        if (yrun) {
            double rdy = 0.0;
            if (j >= 1) rdy = fmax(rdy, 1.0 / s2[j - 1]);
            if (j <= ny - 2) rdy = fmax(rdy, 1.0 / s2[j]);
            cmax = fmax(cmax, (fabs(v) + c) * dt * rdy);
        }
    }
    for (int k = 32; k > 0; k >>= 1) {
        cmax = fmax(cmax, __shfl_xor(cmax, k));
        nbad += __shfl_xor(nbad, k);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&out[0], (unsigned long long)__double_as_longlong(cmax));
        if (nbad) atomicAdd(&out[1], nbad);
    }
}

// Nest feed from the FP64 state (feed_children_grid.f:10-179).
struct Nest64Args {
    const double* cel;
    const double* xv;
    const double* yv;
    const double* dep;
    int nx;
    int nxc, nyc;
    const int* we_ny;  // 1-based parent rows per child row, weights
    const double* we_cy;
    const int* sn_nx;  // 1-based parent columns per child column, weights
    const double* sn_cx;
    int w_nx, e_nx, s_ny, n_ny;
    double w_cx, e_cx, s_cy, n_cy;
    int ii1, ii2, kk1, kk2;  // footprint, 1-based inclusive
    double cuke, hmin, thr;
    unsigned* flag;
    double* uvq;  // west (3, nyc), east, south (3, nxc), north
};

// arrival (feed_children_grid.f:26-46): any footprint cell with dep > hmin
// and |c^2/g - dep| > 0.1 cuke or (|u| + |v|) c > cuke; only 1s are stored
__global__ void nest64_arrival_kernel(const Nest64Args a) {
    const int i = a.ii1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int k = a.kk1 + blockIdx.y;
    if (i > a.ii2 || k > a.kk2) return;
    const size_t o = (size_t)(i - 1) + (size_t)a.nx * (k - 1);
    if (!(a.dep[o] > a.hmin)) return;
    const double c = a.cel[o];
    if (fabs(c * c / kG - a.dep[o]) > a.thr || (fabs(a.xv[o]) + fabs(a.yv[o])) * c > a.cuke) *a.flag = 1u;
}

// (u, v, eta) of parent cell (i, k), 1-based, times w; 0 on dry cells
__device__ __forceinline__ void feed64_val(const Nest64Args& a, int i, int k, double w, double out[3]) {
    const size_t o = (size_t)(i - 1) + (size_t)a.nx * (k - 1);
    if (a.dep[o] > a.hmin) {
        const double c = a.cel[o];
        out[0] = w * a.xv[o];
        out[1] = w * a.yv[o];
        out[2] = w * (c * c / kG - a.dep[o]);
    } else {
        out[0] = out[1] = out[2] = 0.0;
    }
}

// one thread per child edge node: the blend across the two parent lines on
// either side of the edge, then along the edge (feed_children_grid.f:53-179)
__global__ void nest64_feed_kernel(const Nest64Args a) {
    const int side = blockIdx.y;  // 0 west, 1 east, 2 south, 3 north
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int len = side < 2 ? a.nyc : a.nxc;
    if (q >= len) return;
    double v0[3], v1[3], w0[3], w1[3], cc;
    double* out;
    if (side < 2) {
        const int ii = side == 0 ? a.w_nx : a.e_nx;
        const double c0 = side == 0 ? a.w_cx : a.e_cx;
        const int kk = a.we_ny[q];
        cc = a.we_cy[q];
        feed64_val(a, ii, kk, c0, v0);
        feed64_val(a, ii + 1, kk, 1.0 - c0, v1);
        feed64_val(a, ii, kk + 1, c0, w0);
        feed64_val(a, ii + 1, kk + 1, 1.0 - c0, w1);
        out = a.uvq + (size_t)side * 3 * a.nyc;
    } else {
        const int kk = side == 2 ? a.s_ny : a.n_ny;
        const double c0 = side == 2 ? a.s_cy : a.n_cy;
        const int ii = a.sn_nx[q];
        cc = a.sn_cx[q];
        feed64_val(a, ii, kk, c0, v0);
        feed64_val(a, ii, kk + 1, 1.0 - c0, v1);
        feed64_val(a, ii + 1, kk, c0, w0);
        feed64_val(a, ii + 1, kk + 1, 1.0 - c0, w1);
        out = a.uvq + (size_t)6 * a.nyc + (size_t)(side - 2) * 3 * a.nxc;
    }
    for (int f = 0; f < 3; f++) out[q + (size_t)len * f] = (v0[f] + v1[f]) * cc + (w0[f] + w1[f]) * (1.0 - cc);
}

int fail64(int code, const std::string& msg) { return cliffs::set_error(code, msg.c_str()); }

#define HIPCHK64(call)                                                                      \
    do {                                                                                    \
        const hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) return fail64(CLIFFS_ERR_HIP, std::string(#call) + ": " +     \
                                                                hipGetErrorString(e_));      \
    } while (0)

}  // namespace

struct cliffs64_handle {
    int dev = 0, nx = 0, ny = 0, mx = 0;
    double dt = 0, ground = 0, crough = 0, celmin = 0;
    bool xrun = false, yrun = false;
    int ii1 = 1, ii2 = 1, jj1 = 1, jj2 = 1;  // line ranges, 1-based (cliffs_main.f:140-164)
    bool state = false;
    double *dep = nullptr, *cel = nullptr, *xvel = nullptr, *yvel = nullptr;
    double *hmax = nullptr, *velmax = nullptr;
    double *celT = nullptr, *xvelT = nullptr, *yvelT = nullptr, *depT = nullptr;  // column-major copies
    double *s1 = nullptr, *s2 = nullptr, *zeta = nullptr;
    double* edges = nullptr;  // edge1 (mx, 3) then edge2 (mx, 3)
    unsigned long long* health = nullptr;  // cliffs64_health result: Courant bits, non-finite count
    void* gbuf = nullptr;                   // cliffs64_gauges: samples, then indices
    size_t gcap = 0;
    int users = 0;                          // nests of this parent (destroyed first)
    double* wd = nullptr;
    int* wi = nullptr;
    hipStream_t st = nullptr;
};

struct cliffs64_nest {
    cliffs64_handle* parent = nullptr;
    Nest64Args a{};
    void* dmem = nullptr;  // feed output, maps, arrival flag
};

namespace {

void free64(cliffs64_handle* h) {
    if (!h) return;
    (void)hipSetDevice(h->dev);
    if (h->st) (void)hipStreamSynchronize(h->st);
    for (double* p : {h->dep, h->cel, h->xvel, h->yvel, h->hmax, h->velmax, h->s1, h->s2, h->zeta, h->edges,
                      h->wd, h->celT, h->xvelT, h->yvelT, h->depT})
        if (p) (void)hipFree(p);
    if (h->wi) (void)hipFree(h->wi);
    if (h->health) (void)hipFree(h->health);
    if (h->gbuf) (void)hipFree(h->gbuf);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}

hipError_t transpose64(const double* in, double* out, int rows, int cols, hipStream_t st) {
    hipLaunchKernelGGL(transpose64_kernel, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(256), 0, st, in, out, rows,
                       cols);
    return hipGetLastError();
}

// with_max: a y-sweep also updates the running maxima (the step's last sweep)
int sweep64(cliffs64_handle* h, int iu, int first, int last, const double* e1, const double* e2, int mx,
            bool with_max = false) {
    const int n = iu == 1 ? h->nx : h->ny;
    const int nx = h->nx, ny = h->ny;
    if (iu == 2)  // columns as contiguous lines
        for (int f = 0; f < 3; f++)
            HIPCHK64(transpose64(f == 0 ? h->cel : f == 1 ? h->xvel : h->yvel,
                                 f == 0 ? h->celT : f == 1 ? h->xvelT : h->yvelT, ny, nx, h->st));
    Line64 a;
    a.nx = nx;
    a.iu = iu;
    a.first = first;
    a.nlines = last - first + 1;
    a.n = n;
    a.ld = iu == 1 ? nx : ny;
    a.dep = iu == 1 ? h->dep : h->depT;
    a.cel = iu == 1 ? h->cel : h->celT;
    a.xvel = iu == 1 ? h->xvel : h->xvelT;
    a.yvel = iu == 1 ? h->yvel : h->yvelT;
    a.s1 = h->s1;
    a.s2 = h->s2;
    a.zeta = h->zeta;
    a.e1 = e1;
    a.e2 = e2;
    a.mx = mx;
    a.dt = h->dt;
    a.ground = h->ground;
    a.crough = h->crough;
    a.celmin = h->celmin;
    a.wd = h->wd;
    a.wi = h->wi;
    hipLaunchKernelGGL(line64_kernel, dim3((a.nlines + kLineWaves - 1) / kLineWaves), dim3(64 * kLineWaves), 0,
                       h->st, a);
    HIPCHK64(hipGetLastError());
    if (iu == 2 && with_max) {
        hipLaunchKernelGGL(untranspose_max64_kernel, dim3((ny + 31) / 32, (nx + 31) / 32), dim3(256), 0, h->st,
                           h->celT, h->xvelT, h->yvelT, h->cel, h->xvel, h->yvel, h->hmax, h->velmax, nx, ny);
        HIPCHK64(hipGetLastError());
    } else if (iu == 2) {
        for (int f = 0; f < 3; f++)
            HIPCHK64(transpose64(f == 0 ? h->celT : f == 1 ? h->xvelT : h->yvelT,
                                 f == 0 ? h->cel : f == 1 ? h->xvel : h->yvel, nx, ny, h->st));
    }
    return CLIFFS_OK;
}

int maxwave64(cliffs64_handle* h) {
    const size_t m = (size_t)h->nx * h->ny;
    hipLaunchKernelGGL(maxwave64_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, h->st, h->cel, h->xvel,
                       h->yvel, h->hmax, h->velmax, m);
    HIPCHK64(hipGetLastError());
    return CLIFFS_OK;
}

}  // namespace

extern "C" {

int cliffs64_create(const cliffs_grid_desc* d, cliffs64_handle** out) {
    if (!d || !out) return fail64(CLIFFS_ERR_ARG, "NULL descriptor or output pointer");
    *out = nullptr;
    if (d->nx < 1 || d->ny < 1 || !d->dep) return fail64(CLIFFS_ERR_ARG, "grid must be at least 1x1 with a depth array");
    if (!(d->ground > 0.0)) return fail64(CLIFFS_ERR_ARG, "ground must be > 0 (c_min = sqrt(g ground))");
    if (!(d->dt > 0.0)) return fail64(CLIFFS_ERR_ARG, "dt must be > 0");
    if (d->nx > 2 && !d->s1) return fail64(CLIFFS_ERR_ARG, "s1 required when nx > 2");
    if (d->ny > 2 && (!d->s2 || !d->zeta)) return fail64(CLIFFS_ERR_ARG, "s2 and zeta required when ny > 2");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail64(CLIFFS_ERR_NODEVICE, "no HIP device");
    if (d->device < 0 || d->device >= ndev) return fail64(CLIFFS_ERR_NODEVICE, "device index out of range");

    cliffs64_handle* h = new (std::nothrow) cliffs64_handle;
    if (!h) return fail64(CLIFFS_ERR_NOMEM, "host allocation");
    const int nx = d->nx, ny = d->ny;
    h->dev = d->device;
    h->nx = nx;
    h->ny = ny;
    h->dt = d->dt;
    h->ground = d->ground;
    h->crough = d->crough;
    h->celmin = std::sqrt(kG * d->ground);  // cliffs_main.f:227
    h->xrun = nx > 2;
    h->yrun = ny > 2;
    h->ii1 = h->xrun ? 2 : 1;
    h->ii2 = h->xrun ? nx - 1 : 1;
    h->jj1 = h->yrun ? 2 : 1;
    h->jj2 = h->yrun ? ny - 1 : 1;
    h->mx = std::max(h->xrun ? ny : 1, h->yrun ? nx : 1);

    auto bail = [&](int rc) {
        free64(h);
        return rc;
    };
    if (hipSetDevice(h->dev) != hipSuccess) return bail(fail64(CLIFFS_ERR_HIP, "hipSetDevice"));
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess)
        return bail(fail64(CLIFFS_ERR_HIP, "stream creation"));
    const size_t m = (size_t)nx * ny;
    // work arrays for the longer of the two sweeps' (nodes + 2) x lines
    const size_t wn = std::max((size_t)(nx + 2) * ny, (size_t)(ny + 2) * nx);
    const size_t n_s1 = nx > 1 ? (size_t)(nx - 1) * ny : 1, n_s2 = ny > 1 ? (size_t)(ny - 1) : 1;
    bool ok = hipMalloc(&h->dep, m * 8) == hipSuccess && hipMalloc(&h->cel, m * 8) == hipSuccess &&
              hipMalloc(&h->xvel, m * 8) == hipSuccess && hipMalloc(&h->yvel, m * 8) == hipSuccess &&
              hipMalloc(&h->hmax, m * 8) == hipSuccess && hipMalloc(&h->velmax, m * 8) == hipSuccess &&
              hipMalloc(&h->s1, n_s1 * 8) == hipSuccess && hipMalloc(&h->s2, n_s2 * 8) == hipSuccess &&
              hipMalloc(&h->zeta, (size_t)ny * 8) == hipSuccess &&
              hipMalloc(&h->edges, (size_t)6 * h->mx * 8) == hipSuccess &&
              hipMalloc(&h->wd, (size_t)kStash * std::max(nx, ny) * 8) == hipSuccess &&
              hipMalloc(&h->wi, (size_t)kInts * wn * 4) == hipSuccess &&
              hipMalloc(&h->health, 2 * sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc(&h->celT, m * 8) == hipSuccess && hipMalloc(&h->xvelT, m * 8) == hipSuccess &&
              hipMalloc(&h->yvelT, m * 8) == hipSuccess && hipMalloc(&h->depT, m * 8) == hipSuccess;
    if (!ok) return bail(fail64(CLIFFS_ERR_NOMEM, "device allocation"));
    ok = hipMemcpyAsync(h->dep, d->dep, m * 8, hipMemcpyHostToDevice, h->st) == hipSuccess &&
         (nx <= 2 || hipMemcpyAsync(h->s1, d->s1, n_s1 * 8, hipMemcpyHostToDevice, h->st) == hipSuccess) &&
         (ny <= 2 || (hipMemcpyAsync(h->s2, d->s2, n_s2 * 8, hipMemcpyHostToDevice, h->st) == hipSuccess &&
                      hipMemcpyAsync(h->zeta, d->zeta, (size_t)ny * 8, hipMemcpyHostToDevice, h->st) ==
                          hipSuccess)) &&
         hipMemsetAsync(h->cel, 0, m * 8, h->st) == hipSuccess &&
         hipMemsetAsync(h->xvel, 0, m * 8, h->st) == hipSuccess &&
         hipMemsetAsync(h->yvel, 0, m * 8, h->st) == hipSuccess &&
         hipMemsetAsync(h->hmax, 0, m * 8, h->st) == hipSuccess &&
         hipMemsetAsync(h->velmax, 0, m * 8, h->st) == hipSuccess &&
         transpose64(h->dep, h->depT, ny, nx, h->st) == hipSuccess &&
         hipStreamSynchronize(h->st) == hipSuccess;
    if (!ok) return bail(fail64(CLIFFS_ERR_HIP, "grid upload"));
    *out = h;
    return CLIFFS_OK;
}

int cliffs64_destroy(cliffs64_handle* h) {
    if (h && h->users > 0) return fail64(CLIFFS_ERR_STATE, "grid still used by a nest: destroy the nest first");
    free64(h);
    return CLIFFS_OK;
}

int cliffs64_upload_state(cliffs64_handle* h, const double* cel, const double* xvel, const double* yvel) {
    if (!h || !cel || !xvel || !yvel) return fail64(CLIFFS_ERR_ARG, "NULL argument");
    HIPCHK64(hipSetDevice(h->dev));
    const size_t b = (size_t)h->nx * h->ny * 8;
    HIPCHK64(hipMemcpyAsync(h->cel, cel, b, hipMemcpyHostToDevice, h->st));
    HIPCHK64(hipMemcpyAsync(h->xvel, xvel, b, hipMemcpyHostToDevice, h->st));
    HIPCHK64(hipMemcpyAsync(h->yvel, yvel, b, hipMemcpyHostToDevice, h->st));
    HIPCHK64(hipStreamSynchronize(h->st));
    h->state = true;
    return CLIFFS_OK;
}

int cliffs64_download_state(cliffs64_handle* h, double* cel, double* xvel, double* yvel) {
    if (!h || !cel || !xvel || !yvel) return fail64(CLIFFS_ERR_ARG, "NULL argument");
    HIPCHK64(hipSetDevice(h->dev));
    const size_t b = (size_t)h->nx * h->ny * 8;
    HIPCHK64(hipMemcpyAsync(cel, h->cel, b, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipMemcpyAsync(xvel, h->xvel, b, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipMemcpyAsync(yvel, h->yvel, b, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipStreamSynchronize(h->st));
    return CLIFFS_OK;
}

int cliffs64_download_max(cliffs64_handle* h, double* hmax, double* velmax) {
    if (!h || !hmax || !velmax) return fail64(CLIFFS_ERR_ARG, "NULL argument");
    HIPCHK64(hipSetDevice(h->dev));
    const size_t b = (size_t)h->nx * h->ny * 8;
    HIPCHK64(hipMemcpyAsync(hmax, h->hmax, b, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipMemcpyAsync(velmax, h->velmax, b, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipStreamSynchronize(h->st));
    return CLIFFS_OK;
}

int cliffs64_sweep(cliffs64_handle* h, int iu, int first, int last, const double* edge1, const double* edge2,
                   int mx) {
    if (!h) return fail64(CLIFFS_ERR_ARG, "NULL handle");
    if (!h->state) return fail64(CLIFFS_ERR_STATE, "no state uploaded");
    if (iu != 1 && iu != 2) return fail64(CLIFFS_ERR_ARG, "iu must be 1 or 2");
    const int nl = iu == 1 ? h->ny : h->nx;
    const int n = iu == 1 ? h->nx : h->ny;
    if (first < 1 || last > nl || first > last) return fail64(CLIFFS_ERR_ARG, "line range out of grid");
    if (n < 3) return fail64(CLIFFS_ERR_ARG, "line length must be >= 3");
    if ((edge1 || edge2) && mx < last) return fail64(CLIFFS_ERR_ARG, "edge arrays shorter than the lines swept");
    HIPCHK64(hipSetDevice(h->dev));
    const double *e1 = nullptr, *e2 = nullptr;
    if (edge1 || edge2) {
        if (mx > h->mx) return fail64(CLIFFS_ERR_ARG, "edge arrays longer than the grid's mx");
        const size_t b = (size_t)3 * mx * 8;
        if (edge1) HIPCHK64(hipMemcpyAsync(h->edges, edge1, b, hipMemcpyHostToDevice, h->st));
        if (edge2) HIPCHK64(hipMemcpyAsync(h->edges + (size_t)3 * h->mx, edge2, b, hipMemcpyHostToDevice, h->st));
        e1 = edge1 ? h->edges : nullptr;
        e2 = edge2 ? h->edges + (size_t)3 * h->mx : nullptr;
    }
    int rc = sweep64(h, iu, first, last, e1, e2, mx);
    if (rc) return rc;
    // the edge staging buffer is reused by the next call
    if (e1 || e2) HIPCHK64(hipStreamSynchronize(h->st));
    return CLIFFS_OK;
}

int cliffs64_maxwave(cliffs64_handle* h) {
    if (!h) return fail64(CLIFFS_ERR_ARG, "NULL handle");
    if (!h->state) return fail64(CLIFFS_ERR_STATE, "no state uploaded");
    HIPCHK64(hipSetDevice(h->dev));
    return maxwave64(h);
}

int cliffs64_gauges(cliffs64_handle* h, int n, const int* I, const int* J, double* eta, double* u, double* v) {
    if (!h || n < 0 || (n > 0 && (!I || !J || !eta || !u || !v))) return fail64(CLIFFS_ERR_ARG, "bad gauge arguments");
    if (!h->state) return fail64(CLIFFS_ERR_STATE, "no state uploaded");
    if (n == 0) return CLIFFS_OK;
    for (int q = 0; q < n; q++)
        if (I[q] < 1 || I[q] > h->nx || J[q] < 1 || J[q] > h->ny) return fail64(CLIFFS_ERR_ARG, "gauge outside the grid");
    HIPCHK64(hipSetDevice(h->dev));
    // gauge indices and samples in one device buffer, kept for the next call
    const size_t need = sizeof(double) * 3 * n + sizeof(int) * 2 * n;
    if (need > h->gcap) {
        if (h->gbuf) HIPCHK64(hipFree(h->gbuf));
        h->gbuf = nullptr;
        h->gcap = 0;
        if (hipMalloc(&h->gbuf, need) != hipSuccess) {
            h->gbuf = nullptr;
            return fail64(CLIFFS_ERR_NOMEM, "gauge buffer");
        }
        h->gcap = need;
    }
    double* dout = static_cast<double*>(h->gbuf);
    int* dij = reinterpret_cast<int*>(dout + 3 * (size_t)n);
    HIPCHK64(hipMemcpyAsync(dij, I, sizeof(int) * n, hipMemcpyHostToDevice, h->st));
    HIPCHK64(hipMemcpyAsync(dij + n, J, sizeof(int) * n, hipMemcpyHostToDevice, h->st));
    hipLaunchKernelGGL(gauges64_kernel, dim3((n + 63) / 64), dim3(64), 0, h->st, h->cel, h->xvel, h->yvel, h->dep,
                       h->nx, h->ny, h->celmin, dij, n, dout);
    HIPCHK64(hipGetLastError());
    HIPCHK64(hipMemcpyAsync(eta, dout, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipMemcpyAsync(u, dout + n, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipMemcpyAsync(v, dout + 2 * n, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipStreamSynchronize(h->st));
    return CLIFFS_OK;
}

int cliffs64_health(cliffs64_handle* h, double* courant, long long* nonfinite) {
    if (!h || !courant || !nonfinite) return fail64(CLIFFS_ERR_ARG, "NULL argument");
    if (!h->state) return fail64(CLIFFS_ERR_STATE, "no state uploaded");
    HIPCHK64(hipSetDevice(h->dev));
    HIPCHK64(hipMemsetAsync(h->health, 0, 2 * sizeof(unsigned long long), h->st));
    const size_t m = (size_t)h->nx * h->ny;
    const unsigned blocks = (unsigned)std::min<size_t>((m + 255) / 256, 4096);
    hipLaunchKernelGGL(health64_kernel, dim3(blocks), dim3(256), 0, h->st, h->cel, h->xvel, h->yvel, h->s1, h->s2,
                       h->nx, h->ny, h->xrun ? 1 : 0, h->yrun ? 1 : 0, h->dt, h->celmin, h->health);
    HIPCHK64(hipGetLastError());
    unsigned long long r[2];
    HIPCHK64(hipMemcpyAsync(r, h->health, sizeof(r), hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipStreamSynchronize(h->st));
    double c;
    std::memcpy(&c, &r[0], sizeof(c));
    *courant = c;
    *nonfinite = (long long)r[1];
    return CLIFFS_OK;
}

int cliffs64_run(cliffs64_handle* h, int nsteps) {
    if (!h || nsteps < 0) return fail64(CLIFFS_ERR_ARG, "bad run arguments");
    if (!h->state) return fail64(CLIFFS_ERR_STATE, "no state uploaded");
    HIPCHK64(hipSetDevice(h->dev));
    for (int s = 0; s < nsteps; s++) {
        int rc;
        if (h->xrun && (rc = sweep64(h, 1, h->jj1, h->jj2, nullptr, nullptr, h->mx))) return rc;
        if (h->yrun) {
            if ((rc = sweep64(h, 2, h->ii1, h->ii2, nullptr, nullptr, h->mx, true))) return rc;
        } else if ((rc = maxwave64(h))) {
            return rc;
        }
    }
    HIPCHK64(hipStreamSynchronize(h->st));
    return CLIFFS_OK;
}

int cliffs64_nest_create(cliffs64_handle* parent, int nxc, int nyc, const int* we_ny, const double* we_cy,
                         const int* sn_nx, const double* sn_cx, int w_nx, double w_cx, int e_nx, double e_cx,
                         int s_ny, double s_cy, int n_ny, double n_cy, cliffs64_nest** out) {
    if (!parent || !out || !we_ny || !we_cy || !sn_nx || !sn_cx || nxc < 2 || nyc < 2)
        return fail64(CLIFFS_ERR_ARG, "bad nest arguments");
    *out = nullptr;
    const int nx = parent->nx, ny = parent->ny;
    auto in = [](int v, int lo, int hi) { return v >= lo && v <= hi; };
    for (int q = 0; q < nyc; q++)
        if (!in(we_ny[q], 1, ny - 1)) return fail64(CLIFFS_ERR_ARG, "child row map outside the parent");
    for (int q = 0; q < nxc; q++)
        if (!in(sn_nx[q], 1, nx - 1)) return fail64(CLIFFS_ERR_ARG, "child column map outside the parent");
    if (!in(w_nx, 1, nx - 1) || !in(e_nx, 1, nx - 1) || !in(s_ny, 1, ny - 1) || !in(n_ny, 1, ny - 1))
        return fail64(CLIFFS_ERR_ARG, "child edge lines outside the parent");
    cliffs64_nest* n = new (std::nothrow) cliffs64_nest;
    if (!n) return fail64(CLIFFS_ERR_NOMEM, "host allocation");
    n->parent = parent;
    Nest64Args& a = n->a;
    a.nx = nx;
    a.nxc = nxc;
    a.nyc = nyc;
    a.w_nx = w_nx, a.e_nx = e_nx, a.s_ny = s_ny, a.n_ny = n_ny;
    a.w_cx = w_cx, a.e_cx = e_cx, a.s_cy = s_cy, a.n_cy = n_cy;
    a.kk1 = we_ny[0];
    a.kk2 = we_ny[nyc - 1] + 1;
    a.ii1 = sn_nx[0];
    a.ii2 = sn_nx[nxc - 1] + 1;
    if (hipSetDevice(parent->dev) != hipSuccess ||
        hipMalloc(&n->dmem, sizeof(int) * (nyc + nxc) + sizeof(double) * (nyc + nxc) * 7 + 256) != hipSuccess) {
        delete n;
        return fail64(CLIFFS_ERR_NOMEM, "nest device allocation");
    }
    char* p = static_cast<char*>(n->dmem);
    double* uvq = reinterpret_cast<double*>(p);
    double* wcy = uvq + (size_t)6 * (nyc + nxc);
    double* scx = wcy + nyc;
    int* wny = reinterpret_cast<int*>(scx + nxc);
    int* snx = wny + nyc;
    unsigned* flag = reinterpret_cast<unsigned*>(snx + nxc);
    a.uvq = uvq, a.we_cy = wcy, a.sn_cx = scx, a.we_ny = wny, a.sn_nx = snx, a.flag = flag;
    hipStream_t st = parent->st;
    if (hipMemcpyAsync(wcy, we_cy, sizeof(double) * nyc, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(scx, sn_cx, sizeof(double) * nxc, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(wny, we_ny, sizeof(int) * nyc, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(snx, sn_nx, sizeof(int) * nxc, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        (void)hipFree(n->dmem);
        delete n;
        return fail64(CLIFFS_ERR_HIP, "nest map upload");
    }
    parent->users++;
    *out = n;
    return CLIFFS_OK;
}

int cliffs64_nest_destroy(cliffs64_nest* n) {
    if (!n) return CLIFFS_OK;
    (void)hipSetDevice(n->parent->dev);
    (void)hipStreamSynchronize(n->parent->st);
    (void)hipFree(n->dmem);
    n->parent->users--;
    delete n;
    return CLIFFS_OK;
}

int cliffs64_nest_feed(cliffs64_nest* n, double cuke, double hmin, int* iarv, int* fed, double* uvq_w,
                       double* uvq_e, double* uvq_s, double* uvq_n) {
    if (!n || !iarv || !fed || !uvq_w || !uvq_e || !uvq_s || !uvq_n) return fail64(CLIFFS_ERR_ARG, "NULL argument");
    cliffs64_handle* h = n->parent;
    if (!h->state) return fail64(CLIFFS_ERR_STATE, "no state uploaded");
    HIPCHK64(hipSetDevice(h->dev));
    Nest64Args a = n->a;
    a.cel = h->cel, a.xv = h->xvel, a.yv = h->yvel, a.dep = h->dep;
    a.cuke = cuke, a.hmin = hmin, a.thr = (double)0.1f * cuke;  // feed_children_grid.f:30
    *fed = 0;
    if (!*iarv) {
        HIPCHK64(hipMemsetAsync(a.flag, 0, sizeof(unsigned), h->st));
        const int w = a.ii2 - a.ii1 + 1, ht = a.kk2 - a.kk1 + 1;
        hipLaunchKernelGGL(nest64_arrival_kernel, dim3((w + 255) / 256, ht), dim3(256), 0, h->st, a);
        HIPCHK64(hipGetLastError());
        unsigned f = 0;
        HIPCHK64(hipMemcpyAsync(&f, a.flag, sizeof(f), hipMemcpyDeviceToHost, h->st));
        HIPCHK64(hipStreamSynchronize(h->st));
        if (!f) return CLIFFS_OK;
        *iarv = 1;
    }
    const int len = std::max(a.nxc, a.nyc);
    hipLaunchKernelGGL(nest64_feed_kernel, dim3((len + 127) / 128, 4), dim3(128), 0, h->st, a);
    HIPCHK64(hipGetLastError());
    const size_t bw = sizeof(double) * 3 * a.nyc, bs = sizeof(double) * 3 * a.nxc;
    HIPCHK64(hipMemcpyAsync(uvq_w, a.uvq, bw, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipMemcpyAsync(uvq_e, a.uvq + (size_t)3 * a.nyc, bw, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipMemcpyAsync(uvq_s, a.uvq + (size_t)6 * a.nyc, bs, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipMemcpyAsync(uvq_n, a.uvq + (size_t)6 * a.nyc + (size_t)3 * a.nxc, bs, hipMemcpyDeviceToHost, h->st));
    HIPCHK64(hipStreamSynchronize(h->st));
    *fed = 1;
    return CLIFFS_OK;
}

}  // extern "C"
