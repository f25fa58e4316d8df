// One-way nest feed on the device (feed_children_grid.f:10-179).
//
// arrival_kernel: any wet-enough parent cell of the child footprint with
//   |eta| > 0.1*cuke or (|u|+|v|) c > cuke  (feed_children_grid.f:26-46);
//   dry cells (dep <= hmin, here the device predicate s <= celmin, which
//   equals dep <= ground bit for bit) are skipped.  Any thread that finds
//   one stores 1 into the flag (only 1s are ever stored).
// feed_kernel: one thread per child edge node; the two-point blend across
//   the parent lines on either side of the child edge followed by the
//   two-point blend along the edge (feed_children_grid.f:53-80 and the
//   three mirrored blocks), in FP64 from the FP32 state, dry parent cells
//   contributing 0.
#include <hip/hip_runtime.h>

#include "cliffs_dev.h"

// The FP64 arithmetic of the feed and of the edge interpolation rounds every
// product and sum as the reference's real*8 code does: no FMA contraction.
#pragma clang fp contract(off)

namespace cliffs {

__global__ void nest_arrival_kernel(const NestArgs a) {
    const int i = a.ii1 + blockIdx.x * blockDim.x + threadIdx.x;  // 1-based column
    const int k = a.kk1 + blockIdx.y;                              // 1-based row
    if (i > a.ii2 || k > a.kk2) return;
    const size_t o = (size_t)(k - 1) * a.g.pitch + (i - 1);
    const float s = a.g.S[o];
    if (!is_deep(s, a.p)) return;
    const float cp = a.g.CP[o];
    const float ge = geta_of(s, cp);
    const float c = c0_of(s) + cp;
    const bool hit = fabsf(ge) > a.gthr ||
                     (fabsf(a.g.U[o]) + fabsf(a.g.V[o])) * c > a.cuke;
    if (hit) *a.flag = 1u;
}

// value triple (u, v, eta) of parent cell (i,k), 1-based, times weight w
__device__ __forceinline__ void feed_val(const NestArgs& a, int i, int k, double w, double out[3]) {
    const size_t o = (size_t)(k - 1) * a.g.pitch + (i - 1);
    const float s = a.g.S[o];
    if (is_deep(s, a.p)) {
        out[0] = w * (double)a.g.U[o];
        out[1] = w * (double)a.g.V[o];
        out[2] = w * ((double)geta_of(s, a.g.CP[o]) / a.grav);
    } else {
        out[0] = out[1] = out[2] = 0.0;
    }
}

__global__ void nest_feed_kernel(const NestArgs a) {
    const int side = blockIdx.y;  // 0 west, 1 east, 2 south, 3 north
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int len = side < 2 ? a.nyc : a.nxc;
    if (q >= len) return;
    double v0[3], v1[3], w0[3], w1[3];
    double* out;
    double cc;
    if (side < 2) {
        const int ii = side == 0 ? a.w_nx : a.e_nx;
        const double c0 = side == 0 ? a.w_cx : a.e_cx;
        const int kk = a.we_ny[q];
        cc = a.we_cy[q];
        // clm(kk) and clm(kk+1): blend of parent columns ii and ii+1
        feed_val(a, ii, kk, c0, v0);
        feed_val(a, ii + 1, kk, 1.0 - c0, v1);
        feed_val(a, ii, kk + 1, c0, w0);
        feed_val(a, ii + 1, kk + 1, 1.0 - c0, w1);
        out = a.uvq + (size_t)side * 3 * a.nyc;
    } else {
        const int kk = side == 2 ? a.s_ny : a.n_ny;
        const double c0 = side == 2 ? a.s_cy : a.n_cy;
        const int ii = a.sn_nx[q];
        cc = a.sn_cx[q];
        feed_val(a, ii, kk, c0, v0);
        feed_val(a, ii, kk + 1, 1.0 - c0, v1);
        feed_val(a, ii + 1, kk, c0, w0);
        feed_val(a, ii + 1, kk + 1, 1.0 - c0, w1);
        out = a.uvq + (size_t)6 * a.nyc + (size_t)(side - 2) * 3 * a.nxc;
    }
#pragma unroll
    for (int f = 0; f < 3; f++) out[q + (size_t)len * f] = (v0[f] + v1[f]) * cc + (w0[f] + w1[f]) * (1.0 - cc);
}

// Child edges at time t from two feed records (cliffs_main.f:318-324,
// 333-339): edge = rec0*cc + rec1*(1-cc) in FP64 with every product and sum
// rounded separately (no contraction, as the reference's real*8 code), then
// staged for the sweep kernels exactly as host edges are: FP32 normal
// velocity, tangential velocity and g*eta, [6][n] per sweep direction.
// zero: the feed has ended (edges 0, cliffs_main.f:306-310).
__global__ void child_edges_kernel(const double* r0, const double* r1, double cc, double omc,
                                   double grav, int nxc, int nyc, int zero, float* xe, float* ye) {
    const int side = blockIdx.y;  // 0 west, 1 east, 2 south, 3 north
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int len = side < 2 ? nyc : nxc;
    if (q >= len) return;
    const size_t off = side < 2 ? (size_t)side * 3 * nyc : (size_t)6 * nyc + (size_t)(side - 2) * 3 * nxc;
    double val[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        const size_t o = off + q + (size_t)len * f;
        const double p0 = r0[o] * cc;   // separate roundings (contraction is off)
        const double p1 = r1[o] * omc;
        val[f] = zero ? 0.0 : p0 + p1;
    }
    const int cn = side < 2 ? 0 : 1;  // normal velocity: u for the x-sweep, v for the y-sweep
    float* out = side < 2 ? xe + (size_t)side * 3 * nyc : ye + (size_t)(side - 2) * 3 * nxc;
    out[q] = (float)val[cn];
    out[len + q] = (float)val[1 - cn];
    out[2 * len + q] = (float)(grav * val[2]);
}

// Largest |eta| over the four edges of one feed record (the quantity
// get_boundary_feed.f:147-158 compares with the child's threshold): one
// block of 256 threads; max of absolute values is exact in any order.
__global__ void record_wave_kernel(const double* rec, int nxc, int nyc, double* wave) {
    __shared__ double part[256];
    double m = 0.0;
    const int len = 2 * nyc + 2 * nxc;  // eta entries: W, E (nyc each), S, N (nxc each)
    for (int q = threadIdx.x; q < len; q += 256) {
        size_t o;
        if (q < 2 * nyc) {
            const int side = q / nyc, k = q % nyc;
            o = (size_t)side * 3 * nyc + (size_t)2 * nyc + k;
        } else {
            const int r = q - 2 * nyc, side = r / nxc, k = r % nxc;
            o = (size_t)6 * nyc + (size_t)side * 3 * nxc + (size_t)2 * nxc + k;
        }
        m = fmax(m, fabs(rec[o]));
    }
    part[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *wave = part[0];
}

hipError_t launch_record_wave(const double* rec, int nxc, int nyc, double* wave, hipStream_t st) {
    hipLaunchKernelGGL(record_wave_kernel, dim3(1), dim3(256), 0, st, rec, nxc, nyc, wave);
    return hipGetLastError();
}

hipError_t launch_child_edges(const double* r0, const double* r1, double cc, double omc, double grav,
                              int nxc, int nyc, int zero, float* xe, float* ye, hipStream_t st) {
    const int len = nxc > nyc ? nxc : nyc;
    hipLaunchKernelGGL(child_edges_kernel, dim3((len + 127) / 128, 4), dim3(128), 0, st, r0, r1, cc, omc,
                       grav, nxc, nyc, zero, xe, ye);
    return hipGetLastError();
}

hipError_t launch_nest_arrival(const NestArgs& a, hipStream_t st) {
    const int nxb = a.ii2 - a.ii1 + 1, nyb = a.kk2 - a.kk1 + 1;
    hipLaunchKernelGGL(nest_arrival_kernel, dim3((nxb + 255) / 256, nyb), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_nest_feed(const NestArgs& a, hipStream_t st) {
    const int len = a.nxc > a.nyc ? a.nxc : a.nyc;
    hipLaunchKernelGGL(nest_feed_kernel, dim3((len + 127) / 128, 4), dim3(128), 0, st, a);
    return hipGetLastError();
}

}  // namespace cliffs
