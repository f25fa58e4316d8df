// Learned nest surrogate: offshore gauge series -> max-inundation map and run-up.
//
// A network, trained on runs the solver generates, that maps the surface-
// elevation series of 64 gauges on a nested grid's boundary (512 samples)
// to the grid's 512 x 512 max-inundation map and to its maximum run-up.
// Every layer is a GEMM on the MFMA units (BF16 operands, FP32
// accumulation), written by hand:
//
//   encoder (1-D convolutions along time, all gauges as input channels)
//     E1  kernel 4, stride 4:       (B*128, 4*64)  x (256, 128)
//     E2  kernel 5, stride 1:       (B*132, 5*128) x (640, 128)
//     E3  kernel 5, stride 2:       (B*66, 5*128)  x (640, 256)
//     E4  kernel 5, stride 2:       (B*34, 5*256)  x (1280, 256)
//     pool: maximum and mean over the 32 time rows -> 512 features
//   trunk   D1, D2 dense 512 -> 512
//   heads   HR  dense 512 -> 1                       (run-up)
//           M1  dense 512 -> 4096 (an 8 x 8 x 64 image)
//           T1..T6  2-D transposed convolutions 2 x 2, stride 2:
//                   8^2 x 64 -> 16^2 x 32 -> 32^2 x 16 -> 64^2 x 8 ->
//                   128^2 x 8 -> 256^2 x 4 -> 512^2 x 1
//
// A convolution along time is a GEMM over an overlapping strided view of
// its input, A(m, k) = X[m * stride * cin + k], when each sample's rows are
// stored with zero rows before and after them ("same" padding): no im2col
// copy.  Each sample then has a few compute rows whose window runs into the
// next sample; the GEMM's epilogue drops them and stores the valid rows at
// their place in the next layer's padded layout.  The backward pass forms
// the window gradients dZ W^T with the same GEMM and sums overlapping
// windows back onto the input rows (col2im), masked by the input's ReLU.
// With kernel = stride, a transposed convolution is a GEMM on a reshaped
// activation; the map comes out in quadtree order (8 x 8 base pixel, then
// six 2 x 2 sub-pixel levels) and the host permutes maps to and from it.
// ReLU after every layer but HR and T6.  Loss: w_map * mean squared map
// error + w_runup * mean squared run-up error; Adam on FP32 master weights
// with BF16 copies for the GEMMs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cliffs_surrogate.h"

namespace {

constexpr int kMapSide = 512, kLevels = 6;         // 8 x 8 base pixels, six 2 x 2 levels
constexpr long kMapPix = (long)kMapSide * kMapSide;

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ short to_bf16(float f) {
    const __bf16 b = (__bf16)f;  // round to nearest even
    short s;
    __builtin_memcpy(&s, &b, 2);
    return s;
}
__device__ __forceinline__ float from_bf16(short s) {
    return __uint_as_float(((unsigned)(unsigned short)s) << 16);
}

// ---------------------------------------------------------------------------
// GEMM:  C(row(m), n) = epi( sum_k A(m,k) B(k,n) )
//   A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]  (BF16; any strides,
//   so overlapping convolution windows and the transposed operands of the
//   backward pass need no copies)
//   epi: + bias[n]; ReLU; or mask: C = (mask(row,n) > 0) ? acc : 0 (ReLU
//   backward); accumulate: C += result.  Outputs FP32 C (ldc) and/or BF16 Cb.
//   row(m) = m, or with rp > 0: t = m % rp, rows with t >= rv are dropped
//   and the others stored at row (m / rp) * op + osh + t.
// Tiles: 64 x 64 per 256-thread block, K in stages of 64 through LDS (four
// MFMA k-steps per stage between two barriers: a stage's loads are the
// latency a small product waits on); each wave computes a 32 x 32 quarter
// as 2 x 2 MFMA 16x16x16 tiles.  A product with few output tiles is split
// along K (blockIdx.z) into FP32 partials; ksum_kernel adds them in a fixed
// order and applies the epilogue.
constexpr int kTM = 64, kTN = 64, kTK = 64;   // stage depth kTK, or 16 for K <= 32

struct GemmArgs {
    int M, N, K;
    const short* A;
    long sam, sak;
    const short* B;
    long sbk, sbn;
    float* C;       // FP32 output (or null)
    short* Cb;      // BF16 output (or null)
    long ldc;
    const float* bias;
    const short* mask;  // BF16, same layout as C (or null)
    int relu, accumulate;
    int rp, rv, op, osh;  // output row map (rp = 0: identity)
    int kchunk;           // split K: block z sums k in [z kchunk, (z+1) kchunk) ...
    float* part;          // ... into part[z][M][N] (no epilogue; 0: no split)
};

// the epilogue of output (m, n) with accumulated value v
__device__ __forceinline__ void gemm_store(const GemmArgs& g, int m, int n, float v) {
    long row = m;
    if (g.rp > 0) {
        const int tt = m % g.rp;
        if (tt >= g.rv) return;
        row = (long)(m / g.rp) * g.op + g.osh + tt;
    }
    const long o = row * g.ldc + n;
    if (g.bias) v += g.bias[n];
    if (g.relu) v = fmaxf(v, 0.f);
    if (g.mask && !(from_bf16(g.mask[o]) > 0.f)) v = 0.f;
    if (g.accumulate) v += g.C[o];
    if (g.C) g.C[o] = v;
    if (g.Cb) g.Cb[o] = to_bf16(v);
}

template <int TK>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(const GemmArgs g) {
    constexpr int kPer = TK / 4;         // elements of each operand per thread and stage
    constexpr int kSh = TK == 64 ? 6 : 4;
    __shared__ short As[kTM][TK + 8];    // [m][k]
    __shared__ short Bs[kTN][TK + 8];    // [n][k]
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;  // wave's quarter
    const int kbeg = g.part ? blockIdx.z * g.kchunk : 0;
    const int kend = g.part ? min(g.K, kbeg + g.kchunk) : g.K;
    const bool ak = g.sak == 1 || g.sam != 1;   // A's unit stride along k
    const bool bn = g.sbn == 1 || g.sbk != 1;   // B's unit stride along n
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        // stage A (64 x TK) and B (TK x 64), TK / 4 elements each per
        // thread; consecutive lanes take the operand's unit-stride index
        // (k of a row-major A, m of a transposed one; n of B, or k), so a
        // wave's loads are contiguous
        short ra[kPer], rb[kPer];
#pragma unroll
        for (int r = 0; r < kPer; r++) {
            const int e = t + 256 * r;
            const int mm = ak ? e >> kSh : e & 63, kk = ak ? e & (TK - 1) : e >> 6;
            const int m = m0 + mm, k = k0 + kk;
            ra[r] = (m < g.M && k < kend) ? g.A[(long)m * g.sam + (long)k * g.sak] : (short)0;
            const int nn = bn ? e & 63 : e >> kSh, kb = bn ? e >> 6 : e & (TK - 1);
            const int n = n0 + nn, kB = k0 + kb;
            rb[r] = (n < g.N && kB < kend) ? g.B[(long)kB * g.sbk + (long)n * g.sbn] : (short)0;
        }
#pragma unroll
        for (int r = 0; r < kPer; r++) {
            const int e = t + 256 * r;
            if (ak)
                As[e >> kSh][e & (TK - 1)] = ra[r];
            else
                As[e & 63][e >> 6] = ra[r];
            if (bn)
                Bs[e & 63][e >> 6] = rb[r];
            else
                Bs[e >> kSh][e & (TK - 1)] = rb[r];
        }
        __syncthreads();
        // MFMA 16x16x16 operand layout: lane l holds row (A) / column (B)
        // l % 16 and k = 4 (l / 16) .. 4 (l / 16) + 3 of each 16-k step
        const int r16 = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
        for (int ks = 0; ks < TK; ks += 16) {
            bf16x4 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    a[i][q] = As[wm + 16 * i + r16][ks + kq + q];
                    b[i][q] = Bs[wn + 16 * i + r16][ks + kq + q];
                }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // result layout: lane l holds column l % 16, rows 4 (l / 16) + q
    const int c16 = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int m = m0 + wm + 16 * i + rq + q;
                const int n = n0 + wn + 16 * j + c16;
                if (m >= g.M || n >= g.N) continue;
                if (g.part)
                    g.part[((long)blockIdx.z * g.M + m) * g.N + n] = acc[i][j][q];
                else
                    gemm_store(g, m, n, acc[i][j][q]);
            }
}

// Tall, narrow products (N <= 16, many rows: the decoder's last layers, 2 M
// rows x 4 outputs over K = 4): 256 x 16 output tiles, each wave 64 rows as
// four MFMA 16x16x16 tiles, K in stages of 16.
__global__ __launch_bounds__(256) void gemm_bf16_tall_kernel(const GemmArgs g) {
    __shared__ short As[256][16 + 8];    // [m][k]
    __shared__ short Bs[16][16 + 8];     // [n][k]
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * 256, n0 = blockIdx.x * 16;
    const bool ak = g.sak == 1 || g.sam != 1;
    const bool bn = g.sbn == 1 || g.sbk != 1;
    f32x4 acc[4];
#pragma unroll
    for (int i = 0; i < 4; i++) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        short ra[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int e = t + 256 * r;
            const int mm = ak ? e >> 4 : e & 255, kk = ak ? e & 15 : e >> 8;
            const int m = m0 + mm, k = k0 + kk;
            ra[r] = (m < g.M && k < g.K) ? g.A[(long)m * g.sam + (long)k * g.sak] : (short)0;
        }
        const int nn = bn ? t & 15 : t >> 4, kb = bn ? t >> 4 : t & 15;
        const int n = n0 + nn, kB = k0 + kb;
        const short rbv = (n < g.N && kB < g.K) ? g.B[(long)kB * g.sbk + (long)n * g.sbn] : (short)0;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int e = t + 256 * r;
            if (ak)
                As[e >> 4][e & 15] = ra[r];
            else
                As[e & 255][e >> 8] = ra[r];
        }
        Bs[nn][kb] = rbv;
        __syncthreads();
        const int r16 = lane & 15, kq = (lane >> 4) * 4;
        bf16x4 b;
#pragma unroll
        for (int q = 0; q < 4; q++) b[q] = Bs[r16][kq + q];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            bf16x4 a;
#pragma unroll
            for (int q = 0; q < 4; q++) a[q] = As[64 * w + 16 * i + r16][kq + q];
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i], 0, 0, 0);
        }
        __syncthreads();
    }
    const int c16 = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int m = m0 + 64 * w + 16 * i + rq + q;
            const int n = n0 + c16;
            if (m < g.M && n < g.N) gemm_store(g, m, n, acc[i][q]);
        }
}

// the split-K partials added in a fixed order, then the epilogue: a thread
// per output, or for few outputs a block per output (threads over the
// splits, then a tree)
__global__ void ksum_kernel(const float* part, int splits, const GemmArgs g) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)g.M * g.N) return;
    float v = 0.f;
    for (int z = 0; z < splits; z++) v += part[(long)z * g.M * g.N + i];
    gemm_store(g, (int)(i / g.N), (int)(i % g.N), v);
}
__global__ void ksum_block_kernel(const float* part, int splits, const GemmArgs g) {
    __shared__ float red[256];
    const long i = blockIdx.x, mn = (long)g.M * g.N;
    float v = 0.f;
    for (int z = threadIdx.x; z < splits; z += blockDim.x) v += part[(long)z * mn + i];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) gemm_store(g, (int)(i / g.N), (int)(i % g.N), red[0]);
}

constexpr long kSplitWords = 1L << 22;   // split-K partials (FP32 words)

// A product whose few output tiles would leave most of the device idle
// while each block walks a long K (the weight gradients: 4 x 4 outputs
// summed over 2M rows; the trunk: 32 x 512 outputs over K = 512) is split
// along K when a partial buffer is given, to about 256 k per block.
hipError_t gemm(GemmArgs g, hipStream_t st, float* part = nullptr) {
    const dim3 grid((g.N + kTN - 1) / kTN, (g.M + kTM - 1) / kTM);
    const long tiles = (long)grid.x * grid.y;
    int splits = 1;
    if (part && tiles < 256 && g.K >= 1024) {
        // about one resident round of blocks: more splits only queue more
        // latency-bound stages
        splits = (int)std::min<long>({(g.K + 255) / 256, 1024 / tiles + 1, kSplitWords / ((long)g.M * g.N)});
        splits = std::max(splits, 1);
    }
    if (splits > 1) {
        g.kchunk = (int)(((g.K + splits - 1) / splits + kTK - 1) / kTK * kTK);
        splits = (g.K + g.kchunk - 1) / g.kchunk;
    }
    if (splits > 1) {
        g.part = part;
        hipLaunchKernelGGL(gemm_bf16_kernel<kTK>, dim3(grid.x, grid.y, splits), dim3(256), 0, st, g);
        const long n = (long)g.M * g.N;
        if (n < 16384 && splits >= 32)
            hipLaunchKernelGGL(ksum_block_kernel, dim3((unsigned)n), dim3(256), 0, st, (const float*)part, splits, g);
        else
            hipLaunchKernelGGL(ksum_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                               (const float*)part, splits, g);
    } else if (g.N <= 16 && g.M >= 1024) {
        g.part = nullptr;
        hipLaunchKernelGGL(gemm_bf16_tall_kernel, dim3((g.N + 15) / 16, (g.M + 255) / 256), dim3(256), 0, st, g);
    } else {
        g.part = nullptr;
        if (g.K <= 32)
            hipLaunchKernelGGL(gemm_bf16_kernel<16>, grid, dim3(256), 0, st, g);
        else
            hipLaunchKernelGGL(gemm_bf16_kernel<kTK>, grid, dim3(256), 0, st, g);
    }
    return hipGetLastError();
}

// column sums of an FP32 (M x N) matrix, deterministic: partial sums of
// 256-row chunks, then a second pass over the chunks
__global__ void colsum_partial_kernel(const float* X, int M, int N, float* part) {
    const int n = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r = threadIdx.x >> 6;  // 4 row groups
    const int m0 = blockIdx.y * 256;
    float s = 0.f;
    if (n < N)
        for (int m = m0 + r; m < min(m0 + 256, M); m += 4) s += X[(long)m * N + n];
    __shared__ float red[4][64];
    red[r][threadIdx.x & 63] = s;
    __syncthreads();
    if (r == 0 && n < N)
        part[(long)blockIdx.y * N + n] = red[0][threadIdx.x] + red[1][threadIdx.x] +
                                         red[2][threadIdx.x] + red[3][threadIdx.x];
}
// one block per column: each thread sums a fixed stride of the chunks,
// then a tree over the threads (a fixed order: deterministic)
__global__ void colsum_final_kernel(const float* part, int chunks, int N, float* out) {
    __shared__ float red[256];
    const int n = blockIdx.x;
    float s = 0.f;
    for (int c = threadIdx.x; c < chunks; c += blockDim.x) s += part[(long)c * N + n];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = red[0];
}

__global__ void to_bf16_kernel(const float* x, short* y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// Pool over time: per sample and channel, the maximum (first row holding
// it) and the mean of the T rows of X (B*T x C, BF16) -> feat (B x 2C, BF16:
// maxima then means) and the maxima's rows.
__global__ void pool_kernel(const short* X, int B, int T, int C, short* feat, int* arg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * C) return;
    const int b = i / C, c = i - b * C;
    const short* x = X + (long)b * T * C + c;
    float mx = from_bf16(x[0]), sum = 0.f;
    int am = 0;
    for (int t = 0; t < T; t++) {
        const float v = from_bf16(x[(long)t * C]);
        sum += v;
        if (v > mx) {
            mx = v;
            am = t;
        }
    }
    feat[(long)b * 2 * C + c] = to_bf16(mx);
    feat[(long)b * 2 * C + C + c] = to_bf16(sum / (float)T);
    arg[i] = am;
}

// Pool backward: dfeat (B x 2C, FP32) -> dZ of the pooled layer in its
// compute layout (B*P x C, P >= T rows per sample, rows t >= T zero),
// masked by that layer's ReLU output X (B*T x C): FP32 and BF16.
__global__ void pool_back_kernel(const float* dfeat, const int* arg, const short* X, int B, int T,
                                 int P, int C, float* dz, short* dzb) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)B * P * C) return;
    const int c = (int)(i % C);
    const long r = i / C;
    const int b = (int)(r / P), t = (int)(r - (long)b * P);
    float v = 0.f;
    if (t < T && from_bf16(X[((long)b * T + t) * C + c]) > 0.f) {
        v = dfeat[(long)b * 2 * C + C + c] / (float)T;
        if (arg[b * C + c] == t) v += dfeat[(long)b * 2 * C + c];
    }
    dz[i] = v;
    dzb[i] = to_bf16(v);
}

// col2im of a convolution along time (kernel k, stride s, compute rows pc
// per sample, valid vc): the window gradients dcols (B*pc x k*C, FP32) summed
// onto the input rows, which the previous layer stored at period pin with
// shift sh; the result goes to the previous layer's compute layout (pp rows
// per sample, valid vp), masked by its ReLU output act (input layout).
__global__ void col2im_kernel(const float* dcols, const short* act, int B, int C, int k, int s,
                              int pc, int vc, int pin, int sh, int pp, int vp, float* dz, short* dzb) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)B * pp * C) return;
    const int c = (int)(i % C);
    const long r = i / C;
    const int b = (int)(r / pp), t = (int)(r - (long)b * pp);
    float v = 0.f;
    if (t < vp) {
        const int row = sh + t;  // the input row in the padded layout
        if (from_bf16(act[((long)b * pin + row) * C + c]) > 0.f)
            for (int dt = 0; dt < k; dt++) {
                const int u = row - dt;
                if (u < 0 || u % s) continue;
                const int tc = u / s;
                if (tc >= vc) continue;
                v += dcols[((long)b * pc + tc) * (long)(k * C) + (long)dt * C + c];
            }
    }
    dz[i] = v;
    dzb[i] = to_bf16(v);
}

// map loss: sum of (p - y)^2 per block, dP = 2 w (p - y) / n (FP32 and BF16)
__global__ void mse_kernel(const float* p, const float* y, long n, float w, float* dp, short* dpb,
                           float* part) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.f;
    if (i < n) {
        const float d = p[i] - y[i];
        e2 = d * d;
        const float gd = 2.f * w * d / (float)n;
        dp[i] = gd;
        dpb[i] = to_bf16(gd);
    }
    __shared__ float red[256];
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// run-up loss of a batch (one block): sum of squared errors into *part,
// dr = 2 w (r - y) / B (FP32 and BF16)
__global__ void runup_loss_kernel(const float* r, const float* y, int B, float w, float* dr, short* drb,
                                  float* part) {
    __shared__ float red[256];
    float e2 = 0.f;
    for (int i = threadIdx.x; i < B; i += blockDim.x) {
        const float d = r[i] - y[i];
        e2 += d * d;
        const float gd = 2.f * w * d / (float)B;
        dr[i] = gd;
        drb[i] = to_bf16(gd);
    }
    red[threadIdx.x] = e2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *part = w * red[0];
}

// the batch's loss from the partial sums of mse_kernel (blocks of them)
// and runup_loss_kernel (part[blocks]), summed in a fixed order in FP64:
// w_map * mean squared map error + (already weighted) run-up term
__global__ void loss_final_kernel(const float* part, int blocks, long n, float wmap, int b, float* out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < blocks; i += blockDim.x) s += (double)part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)((double)wmap * red[0] / (double)n + (double)part[blocks] / (double)b);
}

// Adam (FP32 master weights), BF16 copy of the new weights
__global__ void adam_kernel(float* w, short* wb, const float* g, float* m, float* v, long n, float lr,
                            float b1, float b2, float c1, float c2, float eps) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = to_bf16(wi);
}

// acc (+)= p over n elements (the networks' quadtree outputs summed)
__global__ void accumulate_kernel(const float* p, long n, int first, float* acc) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc[i] = first ? p[i] : acc[i] + p[i];
}

// Inundation maps in raster order from the networks' summed flood levels
// (quadtree order, `nets` networks): per pixel the level is the mean, the
// map the ground height where ground <= level, else 0.
__global__ void levels_to_maps_kernel(const float* acc, int b, float inv_nets, const float* ground,
                                      float* maps) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)b * kMapPix) return;
    const long bi = i / kMapPix;
    const int pix = (int)(i - bi * kMapPix);
    const int iy = pix / kMapSide, ix = pix % kMapSide;
    int q = (iy >> kLevels) * 8 + (ix >> kLevels);
    for (int k = 1; k <= kLevels; k++)
        q = q * 4 + 2 * ((iy >> (kLevels - k)) & 1) + ((ix >> (kLevels - k)) & 1);
    const float level = acc[bi * kMapPix + q] * inv_nets;
    const float g = ground[pix];
    maps[i] = g <= level ? g : 0.f;
}

// rows of a (n_rows x row_len) FP32 array picked by index, as BF16 and/or FP32
__global__ void gather_rows_kernel(const float* src, const int* idx, int nrows, long row_len,
                                   short* dstb, float* dstf) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)nrows * row_len) return;
    const int r = (int)(i / row_len);
    const long c = i - (long)r * row_len;
    const float v = src[(long)idx[r] * row_len + c];
    if (dstb) dstb[i] = to_bf16(v);
    if (dstf) dstf[i] = v;
}

thread_local std::string g_err;
int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}
#define CHK(call)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return fail(CLIFFS_SUR_ERR_HIP, std::string(#call) + ": " + \
                                                                  hipGetErrorString(e_));  \
    } while (0)

constexpr int kG = 64, kT = 512, kMap = 512;
constexpr long kIn = (long)kG * kT, kOut = (long)kMap * kMap;

// encoder convolutions along time: kernel, stride, channels in/out, input
// rows per sample (padded), compute rows (pin / stride), valid rows, output
// rows per sample as stored for the next layer, and their shift
struct Conv {
    int k, s, cin, cout, pin, pc, valid, pout, sh;
};
constexpr Conv kE[4] = {{4, 4, 64, 128, 512, 128, 128, 132, 2},
                        {5, 1, 128, 128, 132, 132, 128, 132, 2},
                        {5, 2, 128, 256, 132, 66, 64, 68, 2},
                        {5, 2, 256, 256, 68, 34, 32, 32, 0}};
constexpr int kPoolT = 32, kPoolC = 256, kFeat = 2 * kPoolC;
// dense and transposed-convolution layers: rows per sample, in, out, ReLU
struct Dense {
    int rows, in, out, relu;
};
enum { D1, D2, HR, M1, T1, T2, T3, T4, T5, T6, kND };
constexpr Dense kD[kND] = {{1, 512, 512, 1},    {1, 512, 512, 1},     {1, 512, 1, 0},
                           {1, 512, 4096, 1},   {64, 64, 128, 1},     {256, 32, 64, 1},
                           {1024, 16, 32, 1},   {4096, 8, 32, 1},     {16384, 8, 16, 1},
                           {65536, 4, 4, 0}};
constexpr int kLayers = 4 + kND;

long conv_k(int l) { return (long)kE[l].k * kE[l].cin; }
// elements of a stored encoder activation (input of E(l+1)) per sample, and
// the zero slack after the last sample that the next layer's overlapping
// windows of its garbage rows read
long enc_elems(int l, int b) { return (long)b * kE[l].pout * kE[l].cout; }
long enc_slack(int l) { return l < 3 ? conv_k(l + 1) : 0; }

}  // namespace

struct cliffs_sur {
    int dev = 0, bmax = 0;
    hipStream_t st = nullptr;
    long np = 0;                         // parameter count
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 [np]
    short* wb = nullptr;                 // BF16 copy [np]
    long woff[kLayers], boff[kLayers];   // offsets of W_l (in x out, row-major) and b_l
    short* ax = nullptr;                 // input (BF16)
    short* ae[4] = {};                   // encoder outputs, padded layouts
    short* feat = nullptr;               // pooled features (B x 512)
    int* arg = nullptr;                  // rows of the pooled maxima
    short* ad[kND] = {};                 // outputs of D1, D2, M1, T1..T5 (BF16)
    float* pred = nullptr;               // FP32 map output (bmax * kOut)
    float* rpred = nullptr;              // FP32 run-up output (bmax)
    float* yb = nullptr;                 // map targets of the batch
    float* yr = nullptr;                 // run-up targets of the batch
    float* dF = nullptr;                 // FP32 gradient scratch (largest activation)
    short *dA = nullptr, *dB = nullptr;  // BF16 gradient ping-pong
    float* dC = nullptr;                 // FP32 window gradients (col2im input)
    float* dH = nullptr;                 // FP32 trunk gradient (B x 512)
    short* dHb = nullptr;
    float* dR = nullptr;                 // FP32 run-up gradient (B)
    short* dRb = nullptr;
    float* part = nullptr;               // reduction partials
    float* kpart = nullptr;              // split-K partials (kSplitWords)
    float* hpart = nullptr;              // pinned: a loss read back by cliffs_sur_loss_grad
    int* didx = nullptr;                 // batch rows of every training step
    long nidx = 0;
    float* dloss = nullptr;              // per-step losses
    long nloss = 0;
    float* ground = nullptr;             // map ground heights (raster), for cliffs_sur_predict
    float *dx = nullptr, *dy = nullptr, *dyr = nullptr;  // training data on the device
    long ndata = 0;
    int t = 0;                           // Adam step
};

namespace {

long dense_out(int l, int b) { return (long)b * kD[l].rows * kD[l].out; }

GemmArgs plain_gemm(int M, int N, int K, const short* A, long sam, long sak, const short* B, long sbk,
                    long sbn) {
    GemmArgs a{};
    a.M = M;
    a.N = N;
    a.K = K;
    a.A = A;
    a.sam = sam;
    a.sak = sak;
    a.B = B;
    a.sbk = sbk;
    a.sbn = sbn;
    return a;
}

int forward(cliffs_sur* s, int b) {
    // encoder
    for (int l = 0; l < 4; l++) {
        const Conv& E = kE[l];
        const short* in = l == 0 ? s->ax : s->ae[l - 1];
        GemmArgs a = plain_gemm(b * E.pc, E.cout, E.k * E.cin, in, (long)E.s * E.cin, 1,
                                s->wb + s->woff[l], E.cout, 1);
        a.bias = s->w + s->boff[l];
        a.relu = 1;
        a.Cb = s->ae[l];
        a.ldc = E.cout;
        a.rp = E.pc;
        a.rv = E.valid;
        a.op = E.pout;
        a.osh = E.sh;
        CHK(gemm(a, s->st, s->kpart));
    }
    hipLaunchKernelGGL(pool_kernel, dim3((b * kPoolC + 255) / 256), dim3(256), 0, s->st, s->ae[3], b,
                       kPoolT, kPoolC, s->feat, s->arg);
    CHK(hipGetLastError());
    // trunk, run-up head, decoder
    for (int l = 0; l < kND; l++) {
        const Dense& L = kD[l];
        const short* in = l == D1 ? s->feat : (l == D2 || l == HR || l == M1) ? s->ad[l == D2 ? D1 : D2]
                                                                             : s->ad[l - 1];
        const int j = 4 + l;
        GemmArgs a = plain_gemm(b * L.rows, L.out, L.in, in, L.in, 1, s->wb + s->woff[j], L.out, 1);
        a.bias = s->w + s->boff[j];
        a.relu = L.relu;
        a.ldc = L.out;
        if (l == HR)
            a.C = s->rpred;
        else if (l == T6)
            a.C = s->pred;
        else
            a.Cb = s->ad[l];
        CHK(gemm(a, s->st, s->kpart));
    }
    return CLIFFS_SUR_OK;
}

int colsum(cliffs_sur* s, const float* X, int M, int N, float* out) {
    const int chunks = (M + 255) / 256;
    hipLaunchKernelGGL(colsum_partial_kernel, dim3((N + 63) / 64, chunks), dim3(256), 0, s->st, X, M,
                       N, s->part);
    CHK(hipGetLastError());
    hipLaunchKernelGGL(colsum_final_kernel, dim3(N), dim3(256), 0, s->st, s->part, chunks, N, out);
    CHK(hipGetLastError());
    return CLIFFS_SUR_OK;
}

// Gradients of a dense layer j (param index) from dZ (FP32 dz, BF16 dzb,
// M x out): bias and weights; the input's gradient dX = dZ W^T, masked by
// `mask` (the input's ReLU output) when given, into dx (FP32) / dxb (BF16),
// accumulated when `acc`.
int dense_back(cliffs_sur* s, int j, int M, int in, int out, const short* X, const float* dz,
               const short* dzb, float* dx, short* dxb, const short* mask, int acc) {
    int rc;
    if ((rc = colsum(s, dz, M, out, s->g + s->boff[j]))) return rc;
    GemmArgs gw = plain_gemm(in, out, M, X, 1, in, dzb, out, 1);
    gw.C = s->g + s->woff[j];
    gw.ldc = out;
    CHK(gemm(gw, s->st, s->kpart));
    if (!dx && !dxb) return CLIFFS_SUR_OK;
    GemmArgs gx = plain_gemm(M, in, out, dzb, out, 1, s->wb + s->woff[j], 1, out);
    gx.C = dx;
    gx.Cb = dxb;
    gx.ldc = in;
    gx.mask = mask;
    gx.accumulate = acc;
    CHK(gemm(gx, s->st, s->kpart));
    return CLIFFS_SUR_OK;
}

// backward from dP (map, FP32 in dF and BF16 in dA) and dR (run-up):
// parameter gradients into s->g
int backward(cliffs_sur* s, int b) {
    int rc;
    short* dcur = s->dA;
    short* dnext = s->dB;
    // decoder T6 .. M1; M1's input gradient goes to the trunk gradient dH
    for (int l = T6; l >= M1; l--) {
        const Dense& L = kD[l];
        const int M = b * L.rows;
        const short* X = l == M1 ? s->ad[D2] : s->ad[l - 1];
        float* dx = l == M1 ? s->dH : s->dF;
        short* dxb = l == M1 ? s->dHb : dnext;
        if ((rc = dense_back(s, 4 + l, M, L.in, L.out, X, s->dF, dcur, dx, dxb, X, 0))) return rc;
        std::swap(dcur, dnext);
    }
    // run-up head, added to dH
    if ((rc = dense_back(s, 4 + HR, b, 512, 1, s->ad[D2], s->dR, s->dRb, s->dH, s->dHb, s->ad[D2], 1)))
        return rc;
    // trunk: dH is dZ of D2 (masked by its ReLU)
    if ((rc = dense_back(s, 4 + D2, b, 512, 512, s->ad[D1], s->dH, s->dHb, s->dF, s->dA, s->ad[D1], 0)))
        return rc;
    if ((rc = dense_back(s, 4 + D1, b, kFeat, 512, s->feat, s->dF, s->dA, s->dH, s->dHb, nullptr, 0)))
        return rc;
    // pool -> dZ of E4 in its compute layout
    {
        const Conv& E = kE[3];
        const long n = (long)b * E.pc * E.cout;
        hipLaunchKernelGGL(pool_back_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->st, s->dH,
                           s->arg, s->ae[3], b, kPoolT, E.pc, kPoolC, s->dF, s->dA);
        CHK(hipGetLastError());
    }
    // encoder E4 .. E1: dZ in dF / dA
    for (int l = 3; l >= 0; l--) {
        const Conv& E = kE[l];
        const int M = b * E.pc;
        const long K = conv_k(l);
        const short* X = l == 0 ? s->ax : s->ae[l - 1];
        if ((rc = colsum(s, s->dF, M, E.cout, s->g + s->boff[l]))) return rc;
        // dW = A^T dZ over the overlapping windows A(m, k) = X[m s cin + k]
        GemmArgs gw = plain_gemm((int)K, E.cout, M, X, 1, (long)E.s * E.cin, s->dA, E.cout, 1);
        gw.C = s->g + s->woff[l];
        gw.ldc = E.cout;
        CHK(gemm(gw, s->st, s->kpart));
        if (l == 0) break;
        // window gradients dZ W^T, then summed onto the input rows
        GemmArgs gc = plain_gemm(M, (int)K, E.cout, s->dA, E.cout, 1, s->wb + s->woff[l], 1, E.cout);
        gc.C = s->dC;
        gc.ldc = K;
        CHK(gemm(gc, s->st, s->kpart));
        const Conv& P = kE[l - 1];
        const long n = (long)b * P.pc * P.cout;
        hipLaunchKernelGGL(col2im_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->st, s->dC,
                           s->ae[l - 1], b, E.cin, E.k, E.s, E.pc, E.valid, E.pin, P.sh, P.pc, P.valid,
                           s->dF, s->dA);
        CHK(hipGetLastError());
    }
    return CLIFFS_SUR_OK;
}

// loss (into the device word dloss), dP (dF, dA) and dR for the batch in
// ax / yb / yr; nothing waits for the host
int loss_grad(cliffs_sur* s, int b, float wmap, float wrun, float* dloss) {
    const long n = (long)b * kOut;
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(mse_kernel, dim3(blocks), dim3(256), 0, s->st, s->pred, s->yb, n, wmap, s->dF,
                       s->dA, s->part);
    CHK(hipGetLastError());
    hipLaunchKernelGGL(runup_loss_kernel, dim3(1), dim3(256), 0, s->st, s->rpred, s->yr, b, wrun, s->dR,
                       s->dRb, s->part + blocks);
    CHK(hipGetLastError());
    hipLaunchKernelGGL(loss_final_kernel, dim3(1), dim3(256), 0, s->st, (const float*)s->part, blocks, n,
                       wmap, b, dloss);
    CHK(hipGetLastError());
    return CLIFFS_SUR_OK;
}

int upload_batch(cliffs_sur* s, int b, const float* x, const float* y, const float* yr) {
    CHK(hipMemcpyAsync(s->dF, x, sizeof(float) * (size_t)b * kIn, hipMemcpyHostToDevice, s->st));
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)(((long)b * kIn + 255) / 256)), dim3(256), 0,
                       s->st, s->dF, s->ax, (long)b * kIn);
    CHK(hipGetLastError());
    if (y) CHK(hipMemcpyAsync(s->yb, y, sizeof(float) * (size_t)b * kOut, hipMemcpyHostToDevice, s->st));
    if (yr) CHK(hipMemcpyAsync(s->yr, yr, sizeof(float) * (size_t)b, hipMemcpyHostToDevice, s->st));
    return CLIFFS_SUR_OK;
}

void free_sur(cliffs_sur* s) {
    if (!s) return;
    (void)hipSetDevice(s->dev);
    if (s->st) (void)hipStreamSynchronize(s->st);
    for (void* p : {(void*)s->w, (void*)s->g, (void*)s->m, (void*)s->v, (void*)s->wb, (void*)s->ax,
                    (void*)s->feat, (void*)s->arg, (void*)s->pred, (void*)s->rpred, (void*)s->yb,
                    (void*)s->yr, (void*)s->dF, (void*)s->dA, (void*)s->dB, (void*)s->dC, (void*)s->dH,
                    (void*)s->dHb, (void*)s->dR, (void*)s->dRb, (void*)s->part, (void*)s->kpart, (void*)s->didx, (void*)s->dloss,
                    (void*)s->dx, (void*)s->dy, (void*)s->dyr, (void*)s->ground})
        (void)hipFree(p);
    for (auto p : s->ae) (void)hipFree(p);
    for (auto p : s->ad) (void)hipFree(p);
    if (s->hpart) (void)hipHostFree(s->hpart);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

int adam_all(cliffs_sur* s, float lr) {
    s->t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)s->t), c2 = 1.f - std::pow(b2, (float)s->t);
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((s->np + 255) / 256)), dim3(256), 0, s->st, s->w,
                       s->wb, s->g, s->m, s->v, s->np, lr, b1, b2, c1, c2, 1e-8f);
    CHK(hipGetLastError());
    return CLIFFS_SUR_OK;
}

template <class T>
bool dalloc(T** p, size_t n) {
    return hipMalloc((void**)p, sizeof(T) * n) == hipSuccess;
}

}  // namespace

extern "C" {

const char* cliffs_sur_last_error(void) { return g_err.c_str(); }

int cliffs_sur_dims(int* gauges, int* samples, int* map, int* layers) {
    if (gauges) *gauges = kG;
    if (samples) *samples = kT;
    if (map) *map = kMap;
    if (layers) *layers = kLayers;
    return CLIFFS_SUR_OK;
}

int cliffs_sur_create(int device, int batch_max, unsigned seed, cliffs_sur** out) {
    if (!out || batch_max < 1 || batch_max > 256) return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(CLIFFS_SUR_ERR_NODEVICE, "no such HIP device");
    cliffs_sur* s = new (std::nothrow) cliffs_sur;
    if (!s) return fail(CLIFFS_SUR_ERR_NOMEM, "host allocation");
    s->dev = device;
    s->bmax = batch_max;
    auto bail = [&](int rc) {
        free_sur(s);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&s->st) != hipSuccess)
        return bail(fail(CLIFFS_SUR_ERR_HIP, "stream"));
    long np = 0;
    std::vector<long> fan_in(kLayers), nw(kLayers);
    for (int l = 0; l < kLayers; l++) {
        const long in = l < 4 ? conv_k(l) : kD[l - 4].in, o = l < 4 ? kE[l].cout : kD[l - 4].out;
        s->woff[l] = np;
        np += in * o;
        s->boff[l] = np;
        np += o;
        fan_in[l] = in;
        nw[l] = in * o;
    }
    s->np = np;
    const int B = batch_max;
    // FP32 gradient scratch: the largest of the map output, the decoder's
    // activations and the encoder's compute layouts
    long fmax = (long)B * kOut, cmax = 0;
    for (int l = 0; l < kND; l++) fmax = std::max(fmax, dense_out(l, B));
    for (int l = 0; l < 4; l++) {
        fmax = std::max(fmax, (long)B * kE[l].pc * kE[l].cout);
        cmax = std::max(cmax, (long)B * kE[l].pc * conv_k(l));
    }
    bool ok = dalloc(&s->w, np) && dalloc(&s->g, np) && dalloc(&s->m, np) && dalloc(&s->v, np) &&
              dalloc(&s->wb, np) && dalloc(&s->ax, (size_t)B * kIn + conv_k(0)) &&
              dalloc(&s->feat, (size_t)B * kFeat) && dalloc(&s->arg, (size_t)B * kPoolC) &&
              dalloc(&s->pred, (size_t)B * kOut) && dalloc(&s->rpred, (size_t)B) &&
              dalloc(&s->yb, (size_t)B * kOut) && dalloc(&s->yr, (size_t)B) && dalloc(&s->dF, fmax) &&
              dalloc(&s->dA, fmax) && dalloc(&s->dB, fmax) && dalloc(&s->dC, cmax) &&
              dalloc(&s->dH, (size_t)B * kFeat) && dalloc(&s->dHb, (size_t)B * kFeat) &&
              dalloc(&s->dR, (size_t)B) && dalloc(&s->dRb, (size_t)B) &&
              dalloc(&s->part, fmax / 256 + 4096) && dalloc(&s->kpart, kSplitWords) && dalloc(&s->dloss, 1) &&
              hipHostMalloc(&s->hpart, sizeof(float) * 16) == hipSuccess;
    if (ok) s->nloss = 1;
    for (int l = 0; l < 4 && ok; l++) {
        const size_t n = enc_elems(l, B) + enc_slack(l);
        // padding rows and slack are read by overlapping windows: zero once,
        // never written
        ok = dalloc(&s->ae[l], n) && hipMemset(s->ae[l], 0, sizeof(short) * n) == hipSuccess;
    }
    if (ok) ok = hipMemset(s->ax, 0, sizeof(short) * ((size_t)B * kIn + conv_k(0))) == hipSuccess;
    for (int l = 0; l < kND && ok; l++)
        if (l != HR && l != T6) ok = dalloc(&s->ad[l], dense_out(l, B));
    if (!ok) return bail(fail(CLIFFS_SUR_ERR_NOMEM, "device allocation"));
    // He-uniform weights, zero biases (host PRNG: splitmix64); the two
    // linear output layers start 10x smaller
    std::vector<float> hw(np, 0.f);
    unsigned long long z = 0x9E3779B97F4A7C15ull * (seed + 1);
    auto rnd = [&]() {
        z += 0x9E3779B97F4A7C15ull;
        unsigned long long x = z;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < kLayers; l++) {
        const bool last = l == 4 + HR || l == 4 + T6;
        const double lim = std::sqrt(6.0 / (double)fan_in[l]) * (last ? 0.1 : 1.0);
        for (long i = 0; i < nw[l]; i++) hw[s->woff[l] + i] = (float)((2.0 * rnd() - 1.0) * lim);
    }
    const int rc = cliffs_sur_set_params(s, hw.data());
    if (rc) return bail(rc);
    *out = s;
    return CLIFFS_SUR_OK;
}

int cliffs_sur_destroy(cliffs_sur* s) {
    free_sur(s);
    return CLIFFS_SUR_OK;
}

int cliffs_sur_nparams(cliffs_sur* s, long* n) {
    if (!s || !n) return fail(CLIFFS_SUR_ERR_ARG, "NULL argument");
    *n = s->np;
    return CLIFFS_SUR_OK;
}

int cliffs_sur_set_params(cliffs_sur* s, const float* p) {
    if (!s || !p) return fail(CLIFFS_SUR_ERR_ARG, "NULL argument");
    CHK(hipSetDevice(s->dev));
    CHK(hipMemcpy(s->w, p, sizeof(float) * s->np, hipMemcpyHostToDevice));
    CHK(hipMemset(s->m, 0, sizeof(float) * s->np));
    CHK(hipMemset(s->v, 0, sizeof(float) * s->np));
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((s->np + 255) / 256)), dim3(256), 0, s->st, s->w,
                       s->wb, s->np);
    CHK(hipGetLastError());
    CHK(hipStreamSynchronize(s->st));
    s->t = 0;
    return CLIFFS_SUR_OK;
}

int cliffs_sur_get_params(cliffs_sur* s, float* p) {
    if (!s || !p) return fail(CLIFFS_SUR_ERR_ARG, "NULL argument");
    CHK(hipSetDevice(s->dev));
    CHK(hipStreamSynchronize(s->st));
    CHK(hipMemcpy(p, s->w, sizeof(float) * s->np, hipMemcpyDeviceToHost));
    return CLIFFS_SUR_OK;
}

int cliffs_sur_forward(cliffs_sur* s, int b, const float* x, float* y, float* runup) {
    if (!s || !x || (!y && !runup) || b < 1 || b > s->bmax) return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
    CHK(hipSetDevice(s->dev));
    int rc;
    if ((rc = upload_batch(s, b, x, nullptr, nullptr)) || (rc = forward(s, b))) return rc;
    if (y) CHK(hipMemcpyAsync(y, s->pred, sizeof(float) * (size_t)b * kOut, hipMemcpyDeviceToHost, s->st));
    if (runup) CHK(hipMemcpyAsync(runup, s->rpred, sizeof(float) * (size_t)b, hipMemcpyDeviceToHost, s->st));
    CHK(hipStreamSynchronize(s->st));
    return CLIFFS_SUR_OK;
}

int cliffs_sur_loss_grad(cliffs_sur* s, int b, const float* x, const float* y, const float* yr, float wmap,
                         float wrun, float* loss, float* grads) {
    if (!s || !x || !y || !yr || !loss || b < 1 || b > s->bmax)
        return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
    CHK(hipSetDevice(s->dev));
    int rc;
    if ((rc = upload_batch(s, b, x, y, yr)) || (rc = forward(s, b)) ||
        (rc = loss_grad(s, b, wmap, wrun, s->dloss)) || (rc = backward(s, b)))
        return rc;
    CHK(hipMemcpyAsync(s->hpart, s->dloss, sizeof(float), hipMemcpyDeviceToHost, s->st));
    if (grads) CHK(hipMemcpyAsync(grads, s->g, sizeof(float) * s->np, hipMemcpyDeviceToHost, s->st));
    CHK(hipStreamSynchronize(s->st));
    *loss = s->hpart[0];
    return CLIFFS_SUR_OK;
}

int cliffs_sur_set_data(cliffs_sur* s, long n, const float* x, const float* y, const float* yr) {
    if (!s || n < 1 || !x || !y || !yr) return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
    CHK(hipSetDevice(s->dev));
    CHK(hipStreamSynchronize(s->st));
    if (n != s->ndata) {
        (void)hipFree(s->dx);
        (void)hipFree(s->dy);
        (void)hipFree(s->dyr);
        s->dx = s->dy = s->dyr = nullptr;
        s->ndata = 0;
        if (!dalloc(&s->dx, (size_t)n * kIn) || !dalloc(&s->dy, (size_t)n * kOut) || !dalloc(&s->dyr, (size_t)n))
            return fail(CLIFFS_SUR_ERR_NOMEM, "training data allocation");
        s->ndata = n;
    }
    CHK(hipMemcpy(s->dx, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice));
    CHK(hipMemcpy(s->dy, y, sizeof(float) * n * kOut, hipMemcpyHostToDevice));
    CHK(hipMemcpy(s->dyr, yr, sizeof(float) * n, hipMemcpyHostToDevice));
    return CLIFFS_SUR_OK;
}

int cliffs_sur_train(cliffs_sur* s, int b, int steps, float lr, float wmap, float wrun, unsigned seed,
                     float* losses) {
    if (!s || b < 1 || b > s->bmax || steps < 0) return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
    if (s->ndata < 1) return fail(CLIFFS_SUR_ERR_ARG, "no training data (cliffs_sur_set_data)");
    CHK(hipSetDevice(s->dev));
    const long n = s->ndata;
    if (steps == 0) return CLIFFS_SUR_OK;
    // the batches of every step drawn beforehand (epochs of a shuffled
    // order, Fisher-Yates) and copied once, and the losses kept on the
    // device: the steps are queued without waiting for the host
    unsigned long long z = 0xD1B54A32D192ED03ull * (seed + 7);
    std::vector<int> perm(n), idx((size_t)steps * b);
    for (long i = 0; i < n; i++) perm[i] = (int)i;
    long pos = n;
    for (size_t k = 0; k < idx.size(); k++) {
        if (pos >= n) {
            for (long i = n - 1; i > 0; i--) {
                z = z * 6364136223846793005ull + 1442695040888963407ull;
                const long j = (long)((z >> 33) % (unsigned long long)(i + 1));
                std::swap(perm[i], perm[j]);
            }
            pos = 0;
        }
        idx[k] = perm[pos++];
    }
    CHK(hipStreamSynchronize(s->st));
    if ((long)idx.size() > s->nidx) {
        (void)hipFree(s->didx);
        s->didx = nullptr;
        s->nidx = 0;
        if (!dalloc(&s->didx, idx.size())) return fail(CLIFFS_SUR_ERR_NOMEM, "batch index allocation");
        s->nidx = (long)idx.size();
    }
    if (steps > s->nloss) {
        (void)hipFree(s->dloss);
        s->dloss = nullptr;
        s->nloss = 0;
        if (!dalloc(&s->dloss, (size_t)steps)) return fail(CLIFFS_SUR_ERR_NOMEM, "loss allocation");
        s->nloss = steps;
    }
    CHK(hipMemcpyAsync(s->didx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, s->st));
    int rc;
    for (int it = 0; it < steps; it++) {
        const int* bi = s->didx + (size_t)it * b;
        hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)(((long)b * kIn + 255) / 256)), dim3(256), 0,
                           s->st, s->dx, bi, b, kIn, s->ax, (float*)nullptr);
        CHK(hipGetLastError());
        hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)(((long)b * kOut + 255) / 256)), dim3(256), 0,
                           s->st, s->dy, bi, b, kOut, (short*)nullptr, s->yb);
        CHK(hipGetLastError());
        hipLaunchKernelGGL(gather_rows_kernel, dim3(1), dim3(256), 0, s->st, s->dyr, bi, b, 1L,
                           (short*)nullptr, s->yr);
        CHK(hipGetLastError());
        if ((rc = forward(s, b)) || (rc = loss_grad(s, b, wmap, wrun, s->dloss + it)) ||
            (rc = backward(s, b)) || (rc = adam_all(s, lr)))
            return rc;
    }
    if (losses) CHK(hipMemcpyAsync(losses, s->dloss, sizeof(float) * steps, hipMemcpyDeviceToHost, s->st));
    CHK(hipStreamSynchronize(s->st));
    return CLIFFS_SUR_OK;
}

int cliffs_sur_set_ground(cliffs_sur* s, const float* ground) {
    if (!s || !ground) return fail(CLIFFS_SUR_ERR_ARG, "NULL argument");
    CHK(hipSetDevice(s->dev));
    if (!s->ground && !dalloc(&s->ground, kMapPix)) return fail(CLIFFS_SUR_ERR_NOMEM, "ground allocation");
    CHK(hipMemcpyAsync(s->ground, ground, sizeof(float) * kMapPix, hipMemcpyHostToDevice, s->st));
    CHK(hipStreamSynchronize(s->st));
    return CLIFFS_SUR_OK;
}

int cliffs_sur_predict(cliffs_sur* const* nets, int k, int b, const float* x, float* runups, float* maps) {
    if (!nets || k < 1 || !x || b < 1 || (!runups && !maps)) return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
    cliffs_sur* s0 = nets[0];
    for (int i = 0; i < k; i++) {
        if (!nets[i] || nets[i]->dev != s0->dev) return fail(CLIFFS_SUR_ERR_ARG, "networks on different devices");
        if (b > nets[i]->bmax) return fail(CLIFFS_SUR_ERR_ARG, "batch larger than a network's batch_max");
    }
    if (maps && !s0->ground) return fail(CLIFFS_SUR_ERR_ARG, "no ground set (cliffs_sur_set_ground)");
    CHK(hipSetDevice(s0->dev));
    const long n = (long)b * kOut;
    int rc;
    for (int i = 0; i < k; i++) {
        cliffs_sur* s = nets[i];
        if ((rc = upload_batch(s, b, x, nullptr, nullptr)) || (rc = forward(s, b))) return rc;
        if (runups)
            CHK(hipMemcpyAsync(runups + (size_t)i * b, s->rpred, sizeof(float) * b, hipMemcpyDeviceToHost, s->st));
        CHK(hipStreamSynchronize(s->st));
        if (maps) {  // the first network's FP32 scratch holds the sum (its input staging is done)
            hipLaunchKernelGGL(accumulate_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s0->st,
                               (const float*)s->pred, n, i == 0 ? 1 : 0, s0->dF);
            CHK(hipGetLastError());
            CHK(hipStreamSynchronize(s0->st));
        }
    }
    if (maps) {
        hipLaunchKernelGGL(levels_to_maps_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s0->st,
                           (const float*)s0->dF, b, 1.0f / (float)k, (const float*)s0->ground, s0->pred);
        CHK(hipGetLastError());
        CHK(hipMemcpyAsync(maps, s0->pred, sizeof(float) * n, hipMemcpyDeviceToHost, s0->st));
        CHK(hipStreamSynchronize(s0->st));
    }
    return CLIFFS_SUR_OK;
}

int cliffs_sur_gemm_bf16(int device, int M, int N, int K, const float* A, long sam, long sak,
                         const float* B, long sbk, long sbn, float* C) {
    if (!A || !B || !C || M < 1 || N < 1 || K < 1) return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
    CHK(hipSetDevice(device));
    const long na = (long)(M - 1) * sam + (long)(K - 1) * sak + 1;
    const long nb = (long)(K - 1) * sbk + (long)(N - 1) * sbn + 1;
    float *fa = nullptr, *fb = nullptr, *fc = nullptr;
    short *ba = nullptr, *bb = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(fa);
        (void)hipFree(fb);
        (void)hipFree(fc);
        (void)hipFree(ba);
        (void)hipFree(bb);
    };
    if (!dalloc(&fa, na) || !dalloc(&fb, nb) || !dalloc(&fc, (size_t)M * N) || !dalloc(&ba, na) ||
        !dalloc(&bb, nb)) {
        cleanup();
        return fail(CLIFFS_SUR_ERR_NOMEM, "device allocation");
    }
    hipError_t e = hipMemcpy(fa, A, sizeof(float) * na, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(fb, B, sizeof(float) * nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, nullptr, fa, ba, na);
        hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, nullptr, fb, bb, nb);
        GemmArgs g = plain_gemm(M, N, K, ba, sam, sak, bb, sbk, sbn);
        g.C = fc;
        g.ldc = N;
        e = gemm(g, nullptr);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(C, fc, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return fail(CLIFFS_SUR_ERR_HIP, std::string("gemm: ") + hipGetErrorString(e));
    return CLIFFS_SUR_OK;
}

}  // extern "C"
