// Fused VTCS-2 time step for CDNA4 (gfx950): the x-sweep over rows, the
// y-sweep over columns and the max-wave update of one step
// (cliffs_main.f:317-355) in a single pass over HBM.
//
// The reference's y-sweep must see the complete output of the x-sweep, but
// both sweeps are local (stencil radius 3 along the line, plus one flag per
// line), so a tile of output cells needs only its own x-swept rows and
// three halo rows on each side.  A tile is kFStrip = 248 columns by `rows`
// rows; its 256-thread block
//   X: x-sweeps four rows at a time, one per wave (x_row: 256-node window,
//      4 nodes per lane), into a double-buffered LDS ring of four rows;
//   Y: feeds those four rows to the column pipelines (YPipe: one lane per
//      column) and stores the final state and the max-wave fields.
// HBM sees each cell's state read once (plus 3.2 % x-halo and 6 halo rows
// per tile) and written once: 28 B of state plus <= 16 B of max-wave per
// cell-update, instead of 72 B for two separate sweep kernels.
//
// Line flags.  A row is x-swept iff it has a wet node (cliffs.f:113); the
// flag comes from the previous step (rfl, per strip).  A column is y-swept
// iff it has a wet node after the x-sweep of the whole grid, which a tile
// cannot know.  The step kernel predicts it from the state entering the
// step (cpr, per chunk: a column changes that status only when a whole
// column floods or dries) and records the true flag (ctr).  The fixup
// kernel then recomputes, from the unchanged input buffer, every tile
// holding a column whose prediction was wrong, with the true flags; the
// result is identical to the unfused sweeps.  For columns predicted dry the
// max-wave update is deferred to the fixup kernel, so a wrong prediction
// never leaves a wrong maximum behind.  (Sweeping an all-dry column only
// resets its nodes to c = u = v = 0, which cannot raise a maximum, so the
// opposite mistake needs no deferral.)  The step kernel reads those
// columns' maxima and flags the tile (mpend) only if the update would change
// one, so the fixup reads dry land again only where it must.
#include <hip/hip_runtime.h>

#include <vector>

#include "sweep_core.h"

namespace cliffs {

// Column i was wet within some chunk (cpr row i: cstride bytes, one per chunk,
// zero beyond nchunks), read with 16-byte loads that are all in flight at once.
__device__ __forceinline__ bool column_any(const unsigned char* cpr, int cstride, int col) {
    const uint4* q = reinterpret_cast<const uint4*>(cpr + (size_t)col * cstride);
    unsigned any = 0u;
    for (int k = 0; k < cstride / 16; k++) {
        const uint4 v = q[k];
        any |= v.x | v.y | v.z | v.w;
    }
    return any != 0u;
}

// Tiles are assigned to XCDs in contiguous ranges of (chunk, strip) order, so
// neighbouring tiles -- which read each other's halo lines and share the
// 128-byte lines at their borders -- run under the same L2.  Workgroups are
// dispatched round-robin over the XCDs: block b runs on XCD b % kXCDs.
__device__ __forceinline__ int tile_of_block(int b, int ntiles) {
    const int x = b % kXCDs, k = b / kXCDs;
    const int lo = (int)(((long)ntiles * x) / kXCDs);
    const int hi = (int)(((long)ntiles * (x + 1)) / kXCDs);
    return lo + k < hi ? lo + k : -1;
}

template <int N>
struct FShared {
    float row[2][4][4][64 * N];   // [buffer][row of the group][s, cp, u, v][window node]
    float4 xmt[64 * N];           // x metric of the tile's window nodes
    // row metrics of the tile's stream rows, staged once per tile so that no
    // row waits on a global load: y metric and zeta of rows j0-5 .. j1, x row
    // factors of rows j0-3 .. j1+2
    float4 ymt[kFMaxRows + 6];
    float zt[kFMaxRows + 6];
    float2 xrt[kFMaxRows + 6];
    // [wave][row of the tile]: the wave's columns of the row have a wet node
    // after the step.  Each wave writes each row of the tile exactly once
    // (all lanes, the same value): no clearing, no atomics.
    unsigned char rowany[N][kFMaxRows];
};

// Field access by 32-bit element offset from the common base: global
// accesses with a scalar base and a 32-bit vector offset, no 64-bit address
// arithmetic per access.
struct Fields {
    char* base;
    __device__ __forceinline__ float ld(unsigned o) const {
        return *reinterpret_cast<const float*>(base + o * 4u);
    }
    template <int N>
    __device__ __forceinline__ QN<N> ldq(unsigned o) const {
        return cliffs::ldq<N>(reinterpret_cast<const float*>(base + o * 4u));
    }
    __device__ __forceinline__ void st(unsigned o, float v) const {
        *reinterpret_cast<float*>(base + o * 4u) = v;
    }
};

// One tile of the step.  kExact = false: predicted column flags (step
// kernel); kExact = true: the true flags recorded by the step kernel (fixup).
//
// The block's eight waves have two roles.  Waves 0-3 (the X role) x-sweep
// the tile's stream rows four at a time, one row per wave, into the LDS
// ring; waves 4-7 (the Y role, one thread per column) run the column
// pipelines over the rows of the previous group, store the final state and
// the max-wave fields.  Both roles meet at one barrier per group:
//   iteration k:  X: rows of group k -> ring[k & 1]      Y: group k-1 from ring[(k-1) & 1]
// so a role's registers are live only in its own waves, and the x-sweep of
// one group overlaps the y-sweep of the group before.
template <bool kExact, int N>
__device__ __forceinline__ void fused_tile(const FArgs& a, int tile, FShared<N>& sh) {
    constexpr int kW = strip_of(N), kWin = 64 * N, kThreads = fused_threads(N);
    const Params& p = a.p;
    const int nx = a.g.nx, ny = a.g.ny;
    const unsigned pitch = (unsigned)a.g.pitch;
    const Fields F{reinterpret_cast<char*>(a.g.S)};
    const unsigned fs = a.fs;
    const unsigned oCP = (1u + 3u * a.cur) * fs, oU = oCP + fs, oV = oCP + 2u * fs;
    const unsigned oCPn = (4u - 3u * a.cur) * fs, oUn = oCPn + fs, oVn = oCPn + 2u * fs;
    const unsigned oHM = 7u * fs, oVM = 8u * fs;
    const int strip = tile % a.nstrips, chunk = tile / a.nstrips;
    const int cs = strip * kW;
    const int j0 = chunk * a.rows, j1 = min(j0 + a.rows, ny);
    const int t = threadIdx.x;
    const bool xrole = t < kFRoleThreads;               // wave-uniform
    const int lane = t & 63;
    const int w = xrole ? (t >> 6) : ((t - kFRoleThreads) >> 6);  // wave within its role
    const int ty = t - kFRoleThreads;                   // column thread of the Y role

    // staged tables: x metric of the window nodes, row metrics of the stream rows
    for (int k = t; k < kWin; k += kThreads) {
        const int i = cs - 4 + k;
        sh.xmt[k] = (i >= 0 && i < nx) ? a.xm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    const int yb = j0 - 5, xb = j0 - 3;  // first rows of the staged row tables
    for (int k = t; k < j1 - j0 + 6; k += kThreads) {
        const int ry = yb + k, rx = xb + k;
        sh.ymt[k] = (ry >= 0 && ry < ny) ? a.ym[ry] : make_float4(0.f, 0.f, 0.f, 0.f);
        sh.zt[k] = (ry >= 0 && ry < ny) ? a.zeta[ry] : 0.f;
        sh.xrt[k] = (rx >= 0 && rx < ny) ? a.xr[rx] : make_float2(0.f, 0.f);
    }
    __syncthreads();  // xmt and the row tables are complete

    // rows j0-3 .. j1+2 stream through; group g holds stream rows 4g .. 4g+3
    const int nstream = j1 - j0 + 6;
    const int ngroups = (nstream + 3) / 4;

    // X role: lane holds window nodes cs-4+N*lane .. +N-1
    const int i0 = cs - 4 + N * lane;
    const int ld_lane = (xrole && i0 >= 0 && i0 < nx) ? 1 : 0;
    const XLane XL{i0, xlane_vmask<N>(nx, i0), &sh.xmt[N * lane]};

    // Y role: thread ty owns column cs+ty (window node ty+4)
    const int col = cs + ty;
    const int col_ok = (!xrole && ty < kW && col < nx) ? 1 : 0;
    const int in_range = (col_ok && col >= a.col_lo && col <= a.col_hi) ? 1 : 0;
    // max-wave of the column updated here (keepmax) or deferred (loadmax
    // only: the tile is flagged if the update would change a maximum)
    int keepmax = 0, loadmax = 0;
    // bit 0: a node of the column is wet after the step; bit 1: deferred change;
    // bit 2: a node of the column (rows j0 .. j1-1) is wet after the x-sweep
    int colwet_f = 0;

    if (xrole) {
        // ---- X role: the wave's next row is loaded one group ahead
        QN<N> ns = zeroq<N>(), ncp = zeroq<N>(), nu = zeroq<N>(), nv = zeroq<N>();
        int nflag = 0;
        auto x_valid = [&](int g) {
            const int q = j0 - 3 + 4 * g + w;
            return 4 * g + w < nstream && q >= 0 && q < ny;
        };
        auto x_load = [&](int g) {
            if (!x_valid(g)) return;
            const int q = j0 - 3 + 4 * g + w;
            if (ld_lane) {
                const unsigned o = (unsigned)q * pitch + (unsigned)i0;
                ns = F.template ldq<N>(o);
                ncp = F.template ldq<N>(oCP + o);
                nu = F.template ldq<N>(oU + o);
                nv = F.template ldq<N>(oV + o);
            }
            int f = 0;  // row has a wet node in some strip (previous step)
            for (int s0 = 0; s0 < a.nstrips; s0 += 64) {
                const int s = s0 + lane;
                f |= (s < a.nstrips && a.rfl_in[(size_t)s * ny + q] != 0) ? 1 : 0;
            }
            nflag = f;
        };
        x_load(0);
        for (int g = 0; g <= ngroups; g++) {
            // the x rows gate the barrier both roles meet at: the X role
            // runs at raised wave priority against the Y waves on its SIMD
            __builtin_amdgcn_s_setprio(1);
            if (g < ngroups) {
                float(*buf)[4][kWin] = sh.row[g & 1];
                const int q = j0 - 3 + 4 * g + w;
                const bool valid = x_valid(g);
                const QN<N> s = ns, cp = ncp, u = nu, v = nv;
                const bool fl = __ballot(nflag != 0) != 0;
                if (g + 1 < ngroups) x_load(g + 1);
                if (valid) {
                    const bool sweep = q >= a.row_lo && q <= a.row_hi && fl;
                    QN<N> ocp, ou, ov;
                    x_row<N>(XL, p, nx, ny, q, sweep, sh.xrt[q - xb], a.xedges, s, cp, u, v, ocp, ou, ov);
                    stq<N>(&buf[w][0][N * lane], s);
                    stq<N>(&buf[w][1][N * lane], ocp);
                    stq<N>(&buf[w][2][N * lane], ou);
                    stq<N>(&buf[w][3][N * lane], ov);
                }
            }
            __syncthreads();
            __builtin_amdgcn_s_setprio(0);
        }
    } else {
        // ---- Y role
        int yflag = 0;
        if (kExact) {
            yflag = (in_range && a.ctr[col] != 0) ? 1 : 0;
        } else if (in_range) {
            yflag = column_any(a.cpr_in, a.cstride, col) ? 1 : 0;
        }
        keepmax = (a.maxwave && col_ok && (kExact || !in_range || yflag)) ? 1 : 0;
        loadmax = (a.maxwave && col_ok) ? 1 : 0;
        YCol cc;
        cc.swept = in_range && yflag;
        cc.load_edges(a.yedges, nx, col, col_ok);
        YPipe pipe;
        pipe.init();
        // max-wave fields of the rows group g finalises (rows j0-6+4g .. +3),
        // loaded one group ahead
        float hc[4], vc[4];
        auto hv_load = [&](int g) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int r = j0 - 6 + 4 * g + k;
                const bool ok = loadmax && r >= j0 && r < j1;
                const unsigned o = (unsigned)max(r, 0) * pitch + (unsigned)col;
                hc[k] = ok ? F.ld(oHM + o) : 0.f;
                vc[k] = ok ? F.ld(oVM + o) : 0.f;
            }
        };
        hv_load(0);
        for (int g = -1; g < ngroups; g++) {
            if (g >= 0) {
                float(*buf)[4][kWin] = sh.row[g & 1];
                float h[4], vm[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    h[k] = hc[k];
                    vm[k] = vc[k];
                }
                if (g + 1 < ngroups) hv_load(g + 1);
                // four rows through the column pipelines
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int idx = 4 * g + k;
                    if (idx < nstream) {
                        const int r = j0 - 6 + idx;
                        const int q = r + 3;
                        const bool rv = col_ok && q >= 0 && q < ny;
                        RawIn rin{0.f, 0.f, 0.f, 0.f};
                        if (rv) {
                            rin.s = buf[k][0][ty + 4];
                            rin.cp = buf[k][1][ty + 4];
                            rin.u = buf[k][3][ty + 4];  // normal velocity of the y-sweep: yvel
                            rin.v = buf[k][2][ty + 4];
                        }
                        const YOut f = pipe.advance(r, rin, rv, cc, ny, sh.ymt, sh.zt, yb, p);
                        // the column's true y-sweep flag (ctr): row q after the x-sweep is
                        // wet (the pipeline's land bit of its newest row; invalid rows are land)
                        if (!kExact && q >= j0 && q < j1 && !pipe.has(YPipe::kLandB)) colwet_f |= 4;
                        if (r >= j0 && r < j1) {
                            bool wet = false;
                            if (col_ok) {
                                const unsigned o = (unsigned)r * pitch + (unsigned)col;
                                F.st(oCPn + o, f.cp);
                                F.st(oUn + o, f.x);
                                F.st(oVn + o, f.y);
                                {  // max-wave (cliffs_main.f:347-355): store on change only
                                    const float sp2 = speed2(f.x, f.y);
                                    const bool ch = f.cp > h[k], cv = sp2 > vm[k];
                                    if (keepmax && (ch || cv)) {
                                        if (ch) F.st(oHM + o, f.cp);
                                        if (cv) F.st(oVM + o, sp2);
                                    }
                                    colwet_f |= (ch || cv) ? 2 : 0;  // read only for deferred columns
                                }
                                wet = !is_land(c0_of(f.s) + f.cp, p);
                                colwet_f |= wet ? 1 : 0;
                            }
                            sh.rowany[w][r - j0] = __ballot(wet) != 0 ? 1 : 0;
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    const int tile_mchange = __syncthreads_or((colwet_f & 2) && loadmax && !keepmax);
    if (!kExact && t == 0) a.mpend[tile] = tile_mchange ? 1 : 0;
    for (int k = t; k < j1 - j0; k += kThreads) {
        unsigned char any = 0;
#pragma unroll
        for (int y = 0; y < N; y++) any |= sh.rowany[y][k];
        a.rfl_out[(size_t)strip * ny + j0 + k] = any ? 1 : 0;
    }
    if (col_ok) a.cpr_out[(size_t)col * a.cstride + chunk] = colwet_f & 1;
    if (!kExact && col_ok && (colwet_f & 4)) a.ctr[col] = 1;
}

// Eight waves per block (two roles) for 248-column tiles, six for 120-column
// tiles; four waves per SIMD (<= 128 VGPRs), each wave's registers holding
// one role's state only.
template <int N>
__global__ __launch_bounds__(fused_threads(N), 4) void step_fused_kernel(const FArgs a) {
    __shared__ FShared<N> sh;
    const int tile = tile_of_block(blockIdx.x, a.ntiles);
    if (tile < 0) return;
    fused_tile<false, N>(a, tile, sh);
}

// Checks the step kernel's column predictions of one tile: recomputes the
// tile if a prediction was wrong, applies the deferred max-wave update of
// columns that stayed dry, and clears the next step's ctr.
template <int N>
__device__ __forceinline__ void fixup_tile(const FArgs& a, int tile, FShared<N>& sh) {
    constexpr int kW = strip_of(N);
    const int nx = a.g.nx, ny = a.g.ny;
    const int strip = tile % a.nstrips, chunk = tile / a.nstrips;
    const int t = threadIdx.x;
    const int col = strip * kW + t;
    const bool col_ok = t < kW && col < nx;
    const bool in_range = col_ok && col >= a.col_lo && col <= a.col_hi;
    const bool pred = in_range && column_any(a.cpr_in, a.cstride, col);
    const bool tru = in_range && a.ctr[col] != 0;
    if (chunk == 0 && col_ok) a.ctr_next[col] = 0;
    if (__syncthreads_or(in_range && pred != tru)) {
        fused_tile<true, N>(a, tile, sh);
        return;
    }
    // columns left unswept: max-wave of their final state, in tiles where the
    // step kernel found it changes a maximum (after the first steps, dry land
    // changes none); store on change only, as the step kernel does
    if (a.maxwave && a.mpend[tile] && in_range && !pred) {
        const int j0 = chunk * a.rows, j1 = min(j0 + a.rows, ny);
        for (int r = j0; r < j1; r++) {
            const size_t o = (size_t)r * a.g.pitch + col;
            const float cp = a.g.CPn[o], u = a.g.Un[o], v = a.g.Vn[o];
            const float sp2 = speed2(u, v);
            if (cp > a.g.HM[o]) a.g.HM[o] = cp;
            if (sp2 > a.g.VM[o]) a.g.VM[o] = sp2;
        }
    }
}

template <int N>
__global__ __launch_bounds__(fused_threads(N)) void step_fixup_kernel(const FArgs a) {
    __shared__ FShared<N> sh;
    const int tile = tile_of_block(blockIdx.x, a.ntiles);
    if (tile < 0) return;
    fixup_tile<N>(a, tile, sh);
}

// Ensembles: one launch steps every member grid (all of one tile width).
// bmap[b] = (member, tile) or (-1, -1); args[member] is that member's
// argument block for this step.
template <int N>
__global__ __launch_bounds__(fused_threads(N), 4) void step_fused_batch_kernel(
    const FArgs* __restrict__ args, const int2* __restrict__ bmap) {
    __shared__ FShared<N> sh;
    const int2 bm = bmap[blockIdx.x];
    if (bm.x < 0) return;
    fused_tile<false, N>(args[bm.x], bm.y, sh);
}

template <int N>
__global__ __launch_bounds__(fused_threads(N)) void step_fixup_batch_kernel(
    const FArgs* __restrict__ args, const int2* __restrict__ bmap) {
    __shared__ FShared<N> sh;
    const int2 bm = bmap[blockIdx.x];
    if (bm.x < 0) return;
    fixup_tile<N>(args[bm.x], bm.y, sh);
}

// Line flags of the current state (written to rfl_out, cpr_out) and cleared
// ctr buffers, before the first fused step after any other operation.
// Row flags: one wave per (strip, row), its lanes across the strip's
// columns (coalesced), the wave's vote the flag.
__global__ void fused_rowflags_kernel(const FArgs a) {
    const long wid = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // uniform in a wave
    const int lane = threadIdx.x & 63;
    if (wid >= (long)a.nstrips * a.g.ny) return;
    const int s = (int)(wid / a.g.ny), j = (int)(wid % a.g.ny);
    const int i1 = min((s + 1) * a.strip_w, a.g.nx);
    bool any = false;
    for (int i = s * a.strip_w + lane; i < i1; i += 64) {
        const size_t o = (size_t)j * a.g.pitch + i;
        any = any | !is_land(c0_of(a.g.S[o]) + a.g.CP[o], a.p);
    }
    const bool w = __ballot(any) != 0;
    if (lane == 0) a.rfl_out[wid] = w ? 1 : 0;
}

// Column flags per chunk (consecutive threads: consecutive columns) and the
// cleared ctr buffers.
__global__ void fused_colflags_kernel(const FArgs a, unsigned char* ctr0, unsigned char* ctr1) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = a.g.nx, ny = a.g.ny;
    if (t < nx) {
        ctr0[t] = 0;
        ctr1[t] = 0;
    }
    if (t < (long)a.nchunks * nx) {
        const int c = (int)(t / nx), i = (int)(t % nx);
        const int j1 = min((c + 1) * a.rows, ny);
        bool any = false;  // no early exit: the rows' loads stay independent
#pragma unroll 8
        for (int j = c * a.rows; j < j1; j++) {
            const size_t o = (size_t)j * a.g.pitch + i;
            any = any | !is_land(c0_of(a.g.S[o]) + a.g.CP[o], a.p);
        }
        a.cpr_out[(size_t)i * a.cstride + c] = any ? 1 : 0;
    }
}

// Surface elevation at every member's gauges (output_gages.f:18-29: eta on
// wet cells, 0 on dry ones) from the state the argument blocks `args` read:
// member m = blockIdx.y, gauge q; 1-based I = ij[m*n + q], J = ij[(nm + m)*n + q];
// series[(m * nsamples + slot) * n + q].
__global__ void ensemble_gauge_kernel(const FArgs* __restrict__ args, int nm, const int* __restrict__ ij,
                                      int n, int nsamples, int slot, float* __restrict__ series) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (q >= n) return;
    const FArgs& a = args[m];
    const int i = ij[(size_t)m * n + q] - 1, j = ij[((size_t)nm + m) * n + q] - 1;
    const size_t o = (size_t)j * a.g.pitch + i;
    const float s = a.g.S[o], cp = a.g.CP[o];
    series[((size_t)m * nsamples + slot) * n + q] = is_land(c0_of(s) + cp, a.p) ? 0.0f : geta_of(s, cp) / a.p.grav;
}

// ---------------------------------------------------------------------------
// launchers
static int fused_blocks(const FArgs& a) { return (a.ntiles + kXCDs - 1) / kXCDs * kXCDs; }

hipError_t launch_ensemble_gauges(const FArgs* args, int nm, const int* ij, int n, int nsamples, int slot,
                                  float* series, hipStream_t st) {
    hipLaunchKernelGGL(ensemble_gauge_kernel, dim3((n + 63) / 64, nm), dim3(64), 0, st, args, nm, ij, n,
                       nsamples, slot, series);
    return hipGetLastError();
}

hipError_t launch_step_fused(const FArgs& a, hipStream_t st) {
    if (a.strip_w == kFStrip)
        hipLaunchKernelGGL(step_fused_kernel<4>, dim3(fused_blocks(a)), dim3(fused_threads(4)), 0, st, a);
    else
        hipLaunchKernelGGL(step_fused_kernel<2>, dim3(fused_blocks(a)), dim3(fused_threads(2)), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_step_fixup(const FArgs& a, hipStream_t st) {
    if (a.strip_w == kFStrip)
        hipLaunchKernelGGL(step_fixup_kernel<4>, dim3(fused_blocks(a)), dim3(fused_threads(4)), 0, st, a);
    else
        hipLaunchKernelGGL(step_fixup_kernel<2>, dim3(fused_blocks(a)), dim3(fused_threads(2)), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_step_batch(const FArgs* args, const int2* bmap, int nblocks, int strip_w, hipStream_t st) {
    if (strip_w == kFStrip)
        hipLaunchKernelGGL(step_fused_batch_kernel<4>, dim3(nblocks), dim3(fused_threads(4)), 0, st, args, bmap);
    else
        hipLaunchKernelGGL(step_fused_batch_kernel<2>, dim3(nblocks), dim3(fused_threads(2)), 0, st, args, bmap);
    return hipGetLastError();
}

hipError_t launch_fixup_batch(const FArgs* args, const int2* bmap, int nblocks, int strip_w, hipStream_t st) {
    if (strip_w == kFStrip)
        hipLaunchKernelGGL(step_fixup_batch_kernel<4>, dim3(nblocks), dim3(fused_threads(4)), 0, st, args, bmap);
    else
        hipLaunchKernelGGL(step_fixup_batch_kernel<2>, dim3(nblocks), dim3(fused_threads(2)), 0, st, args, bmap);
    return hipGetLastError();
}

hipError_t launch_fused_flags(const FArgs& a, unsigned char* ctr0, unsigned char* ctr1,
                              hipStream_t st) {
    const long rows = (long)a.nstrips * a.g.ny;  // one wave each
    hipLaunchKernelGGL(fused_rowflags_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, st, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long n = (long)a.nchunks * a.g.nx;
    const long m = n > a.g.nx ? n : a.g.nx;
    hipLaunchKernelGGL(fused_colflags_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, a,
                       ctr0, ctr1);
    return hipGetLastError();
}

hipError_t fused_blocks_per_cu(int strip_w, int* n) {
    if (strip_w == kFStrip)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, step_fused_kernel<4>, fused_threads(4), 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, step_fused_kernel<2>, fused_threads(2), 0);
}

// Block map of an ensemble step: every member's tiles, XCD by XCD.  Member
// m's tiles t in [lo(m,x), lo(m,x+1)) go to XCD x in order, so each XCD runs
// contiguous (chunk, strip) ranges of every member (neighbouring tiles share
// its L2); block b runs entry b / kXCDs of XCD b % kXCDs's list.
void ensemble_block_map(const int* ntiles, int n, std::vector<int2>& bmap) {
    std::vector<std::vector<int2>> per(kXCDs);
    for (int m = 0; m < n; m++)
        for (int x = 0; x < kXCDs; x++) {
            const int lo = (int)(((long)ntiles[m] * x) / kXCDs);
            const int hi = (int)(((long)ntiles[m] * (x + 1)) / kXCDs);
            for (int t = lo; t < hi; t++) per[x].push_back(make_int2(m, t));
        }
    size_t len = 0;
    for (auto& v : per) len = v.size() > len ? v.size() : len;
    bmap.assign(len * kXCDs, make_int2(-1, -1));
    for (int x = 0; x < kXCDs; x++)
        for (size_t k = 0; k < per[x].size(); k++) bmap[k * kXCDs + x] = per[x][k];
}

// Tile geometry: enough tiles to fill every CU with `blocks_per_cu` blocks
// once, chunks of at most kFMaxRows rows and at least 32.
void fused_geometry(int nx, int ny, int strip_w, int cus, int blocks_per_cu, int* nstrips, int* nchunks,
                    int* rows) {
    const int ns = (nx + strip_w - 1) / strip_w;
    const int target = cus * blocks_per_cu;
    int nc = target / ns;
    if (nc < 1) nc = 1;
    int r = (ny + nc - 1) / nc;
    if (r < 32) r = 32;
    if (r > kFMaxRows) r = kFMaxRows;
    if (r > ny) r = ny;
    *nstrips = ns;
    *rows = r;
    *nchunks = (ny + r - 1) / r;
}

}  // namespace cliffs
