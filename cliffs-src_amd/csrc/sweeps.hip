// VTCS-2 directional sweeps for CDNA4 (gfx950), one kernel per sweep.
//
// x-sweep (cliffs(1,j,nXn) for all rows, cliffs_main.f:325-329): one wave64
//   per 248-node strip of kXRows rows, 4 nodes per lane (x_row).
// y-sweep (cliffs(2,i,nYn) for all columns, cliffs_main.f:340-344): one lane
//   per column, kYRows output rows per block with a 3-row halo (YPipe); the
//   max-wave update of cliffs_main.f:347-355 is fused into the store.
// These serve single sweeps over line subsets (cliffs_sweep), 1-D grids and
// the unfused 2-D step; the fused 2-D step is in step_fused.hip.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sweep_core.h"

namespace cliffs {

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
    float4 mt[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        mt[k] = (i0 + k >= 0 && i0 + k < nx) ? a.xm[i0 + k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const XLane L{i0, xlane_vmask(nx, i0), mt};
    const int outmask = out_lane ? L.vmask : 0;
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
        const size_t ro = (size_t)j * a.g.pitch + i0;
        const bool sweep = j >= a.act_lo && j <= a.act_hi && a.fin[j];
        const Q4 s = ns, cp = ncp, u = nu, v = nv;
        if (j < r1) load_row(j + 1);
        Q4 ocp, ou, ov;
        x_row(L, p, nx, ny, j, sweep, sweep ? a.xr[j] : make_float2(0.f, 0.f), a.edges, s, cp, u,
              v, ocp, ou, ov);
        // every visited row is written to the other buffer: swept rows with
        // their new values, the others (boundary rows, dry rows) unchanged
        if (out_lane && ld_lane) {
            st4(a.g.CPn + ro, ocp);
            st4(a.g.Un + ro, ou);
            st4(a.g.Vn + ro, ov);
        }
        // wet bits of the new state for the flags consumed by the next sweep
        const int wet = wet_bits(s, ocp, outmask, p);
        if (a.flag_rows) {
            if (__ballot(wet != 0) != 0 && lane == 0) a.fout[j] = 1;
        } else {
            colwet |= wet;
        }
        if (a.fuse_max && out_lane && ld_lane) {  // fused max-wave (cliffs_main.f:347-355)
            float* HMr = a.g.HM + ro;
            float* VMr = a.g.VM + ro;
            Q4 hm = ld4(HMr), vm = ld4(VMr);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                hm.v[k] = fmaxf(hm.v[k], ocp.v[k]);
                vm.v[k] = fmaxf(vm.v[k], speed2(ou.v[k], ov.v[k]));
            }
            st4(HMr, hm);
            st4(VMr, vm);
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

// Rows are fed in batches of kYBatch; the loads of the next batch (state
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
    const int j0 = blockIdx.y * kYRows;
    const int j1 = min(j0 + kYRows, ny);

    YCol cc;
    cc.swept = col_ok && col >= a.col_lo && col <= a.col_hi && a.fin[col];
    cc.load_edges(a.edges, nx, col, col_ok);
    auto row_ok = [&](int row) { return col_ok && row >= 0 && row < ny; };
    YPipe pipe;
    pipe.init();
    unsigned rowwet = 0u;  // lane l: bit0 -> row j0+l, bit1 -> row j0+64+l
    bool colwet = false;

    // feed rows j0-3 .. j1+2: iteration t feeds row j0-3+t and finalises row j0-6+t
    const int r_first = j0 - 6;
    const int n_it = j1 + 3 - (j0 - 3);
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
        for (int k = 0; k < kYBatch; k++) {
            if (it + k < n_it) {
                const int r = r_first + it + k;
                const YOut f = pipe.advance(r, cur[k], row_ok(r + 3), cc, ny, a.ym, a.zeta, 0, p);
                if (r >= j0 && r < j1) {
                    const size_t o = (size_t)r * a.g.pitch + col;
                    bool wet = false;
                    if (col_ok) {
                        // boundary and dry columns are carried over unchanged
                        a.g.CPn[o] = f.cp;
                        a.g.Un[o] = f.x;
                        a.g.Vn[o] = f.y;
                        if (a.fuse_max) {  // fused max-wave (cliffs_main.f:347-355): store on change only
                            const float sp2 = speed2(f.x, f.y);
                            if (f.cp > hc[k]) a.g.HM[o] = f.cp;
                            if (sp2 > vc[k]) a.g.VM[o] = sp2;
                        }
                        wet = !is_land(c0_of(f.s) + f.cp, p);
                        colwet |= wet;
                    }
                    const bool anyw = __ballot(wet) != 0;  // uniform control flow
                    const int rel = r - j0;
                    if (anyw && (rel & 63) == lane) rowwet |= (rel >= 64) ? 2u : 1u;
                }
            }
        }
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
    g.VM[o] = fmaxf(g.VM[o], speed2(u, v));
}

// surface elevation at gauges (1-based I, J) into row `slot` of series:
// eta = g*eta / g on wet cells (output_gages.f:18-29), 0 on dry ones
__global__ void gauge_sample_kernel(DevGrid g, Params p, const int* I, const int* J, int n, float* series,
                                    int slot) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const size_t o = (size_t)(J[q] - 1) * g.pitch + (I[q] - 1);
    const float s = g.S[o], cp = g.CP[o];
    series[(size_t)slot * n + q] = is_land(c0_of(s) + cp, p) ? 0.0f : geta_of(s, cp) / p.grav;
}

// Health of the state the next step reads: the largest Courant number
// (|u| + c) dt / dx over wet cells, in x where rows are swept and in y where
// columns are (dx the smaller spacing next to the cell), and the number of
// cells whose cp, u or v is not finite.  The reference checks neither.
// Thread = column i, rows strided by gridDim.y; wave reduction by
// butterfly shuffles, then one atomic per wave (Courant >= 0, so its bits
// order as unsigned integers).
__global__ void health_kernel(DevGrid g, Params p, const float4* __restrict__ xm,
                              const float2* __restrict__ xr, const float4* __restrict__ ym, int xrun,
                              int yrun, HealthOut* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float cmax = 0.0f;
    unsigned nbad = 0u;
    if (i < g.nx) {
        const float4 mi = xm[i];
        const float rdx = fmaxf(mi.x, mi.y);
        for (int j = blockIdx.y; j < g.ny; j += gridDim.y) {
            const size_t o = (size_t)j * g.pitch + i;
            const float cp = g.CP[o], u = g.U[o], v = g.V[o];
            if (!(isfinite(cp) && isfinite(u) && isfinite(v))) {
                nbad++;
                continue;
            }
            const float c = c0_of(g.S[o]) + cp;
            if (is_land(c, p)) continue;
            if (xrun) cmax = fmaxf(cmax, (fabsf(u) + c) * p.dt * (rdx * xr[j].x));
            if (yrun) {
                const float4 mj = ym[j];
                cmax = fmaxf(cmax, (fabsf(v) + c) * p.dt * fmaxf(mj.x, mj.y));
            }
        }
    }
    for (int k = 32; k > 0; k >>= 1) {
        cmax = fmaxf(cmax, __shfl_xor(cmax, k));
        nbad += __shfl_xor(nbad, k);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&out->cmax_bits, __float_as_uint(cmax));
        if (nbad) atomicAdd(&out->nonfinite, (unsigned long long)nbad);
    }
}

// Running maxima reset (cliffs_reset_max): hmax to c = 0, i.e. cp = -max(s, 0)
// (the expression of the host conversion, exact in FP32), velmax to 0; the
// row padding beyond nx to 0.
__global__ void reset_max_kernel(DevGrid g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.pitch) return;
    const size_t o = (size_t)j * g.pitch + i;
    const float s = g.S[o];
    g.HM[o] = i < g.nx ? -(s > 0.0f ? s : 0.0f) : 0.0f;
    g.VM[o] = 0.0f;
}

// ---------------------------------------------------------------------------
// launchers
hipError_t launch_reset_max(const DevGrid& g, hipStream_t st) {
    hipLaunchKernelGGL(reset_max_kernel, dim3((g.pitch + 255) / 256, g.ny), dim3(256), 0, st, g);
    return hipGetLastError();
}

hipError_t launch_health(const DevGrid& g, const Params& p, const float4* xm, const float2* xr,
                         const float4* ym, int xrun, int yrun, HealthOut* out, hipStream_t st) {
    const dim3 grid((g.nx + 255) / 256, g.ny < 256 ? g.ny : 256);
    hipLaunchKernelGGL(health_kernel, grid, dim3(256), 0, st, g, p, xm, xr, ym, xrun, yrun, out);
    return hipGetLastError();
}

// Device scratch of cliffs_run_gauges_noalloc: the gauge index table and
// the recorded series, kGaugeScratchWords 4-byte words per GPU, held from
// the library's load to the end of the process.  Every call on a GPU shares
// that GPU's copy (calls that overlap on one GPU race on it).  In host code
// `scratch` names a host-side stand-in, not a device address: kernels and
// copies take the per-device address that gauge_scratch looks up.
constexpr size_t kGaugeScratchWords = (size_t)1 << 20;  // 4 MiB
static __device__ __align__(256) unsigned scratch[kGaugeScratchWords];

size_t gauge_scratch_words() { return kGaugeScratchWords; }

// Device address of `scratch` on device `dev`, which must be current: looked
// up once per device (hipGetSymbolAddress), then cached.
hipError_t gauge_scratch(int dev, void** ptr) {
    constexpr int kMaxDev = 64;
    static std::mutex mu;
    static void* addr[kMaxDev] = {};
    if (dev < 0 || dev >= kMaxDev) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    return hipSuccess;
}

hipError_t launch_gauge_sample(const DevGrid& g, const Params& p, const int* I, const int* J, int n,
                               float* series, int slot, hipStream_t st) {
    hipLaunchKernelGGL(gauge_sample_kernel, dim3((n + 255) / 256), dim3(256), 0, st, g, p, I, J, n, series,
                       slot);
    return hipGetLastError();
}

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
