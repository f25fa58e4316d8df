// Learned nest surrogate: offshore gauge series -> max-inundation map.
//
// A small network, trained on runs the solver generates, that maps the
// surface-elevation series of 64 gauges at a nested grid's boundary (256
// samples) to the grid's 64 x 64 max-inundation map.  Every layer is a GEMM
// on the MFMA units (BF16 operands, FP32 accumulation), written by hand:
//
//   encoder  L1  1-D conv over time, kernel 8 = stride 8:  (B*32, 512) x (512, 128)
//            L2  1-D conv, kernel 4 = stride 4:           (B*8, 512)  x (512, 256)
//            L3  dense:                                  (B, 2048)   x (2048, 512)
//   decoder  L4  dense to an 8 x 8 x 64 image:             (B, 512)    x (512, 4096)
//            L5  2-D transposed conv 2x2, stride 2:      (B*64, 64)  x (64, 128)
//            L6  2-D transposed conv 2x2, stride 2:      (B*256, 32) x (32, 64)
//            L7  2-D transposed conv 2x2, stride 2:      (B*1024, 16) x (16, 4)
//
// With kernel = stride, a convolution is a GEMM on a reshaped activation
// (no im2col copies): the input rows (time-major gauge samples) and every
// intermediate activation are laid out so that each layer's output, read
// with the next layer's row length, is exactly the next layer's input.  The
// 64 x 64 output comes out in quadtree order (8 x 8 base pixel, then three
// 2 x 2 sub-pixel levels); the host permutes maps to and from that order.
// ReLU after every layer but the last; MSE loss; Adam on FP32 master
// weights with BF16 copies for the GEMMs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/cliffs_surrogate.h"

namespace {

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
// GEMM:  C(m,n) = epi( sum_k A(m,k) B(k,n) )
//   A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]  (BF16; any strides,
//   so transposed operands of the backward pass need no copies)
//   epi: + bias[n]; ReLU; or mask: C = (mask(m,n) > 0) ? acc : 0 (ReLU
//   backward); accumulate: C += result.  Outputs FP32 C (ldc) and/or BF16 Cb.
// Tiles: 64 x 64 per 256-thread block, K in steps of 16 through LDS; each
// wave computes a 32 x 32 quarter as 2 x 2 MFMA 16x16x16 tiles.
constexpr int kTM = 64, kTN = 64, kTK = 16;

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
};

__global__ __launch_bounds__(256) void gemm_bf16_kernel(const GemmArgs g) {
    __shared__ short As[kTM][kTK + 4];   // [m][k]
    __shared__ short Bs[kTN][kTK + 4];   // [n][k]
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;  // wave's quarter
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        // stage A (64 x 16) and B (16 x 64): 1024 elements each, 4 per thread
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int e = t + 256 * r;
            const int mm = e >> 4, kk = e & 15;      // A: row-major over k
            const int m = m0 + mm, k = k0 + kk;
            As[mm][kk] = (m < g.M && k < g.K) ? g.A[(long)m * g.sam + (long)k * g.sak] : (short)0;
            const int nn = e & 63, kb = e >> 6;      // B: consecutive n
            const int n = n0 + nn, kB = k0 + kb;
            Bs[nn][kb] = (n < g.N && kB < g.K) ? g.B[(long)kB * g.sbk + (long)n * g.sbn] : (short)0;
        }
        __syncthreads();
        // MFMA 16x16x16 operand layout: lane l holds row (A) / column (B)
        // l % 16 and k = 4 (l / 16) .. 4 (l / 16) + 3
        const int r16 = lane & 15, kq = (lane >> 4) * 4;
        bf16x4 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                a[i][q] = As[wm + 16 * i + r16][kq + q];
                b[i][q] = Bs[wn + 16 * i + r16][kq + q];
            }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
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
                float v = acc[i][j][q];
                const long o = (long)m * g.ldc + n;
                if (g.bias) v += g.bias[n];
                if (g.relu) v = fmaxf(v, 0.f);
                if (g.mask && !(from_bf16(g.mask[o]) > 0.f)) v = 0.f;
                if (g.accumulate) v += g.C[o];
                if (g.C) g.C[o] = v;
                if (g.Cb) g.Cb[o] = to_bf16(v);
            }
}

hipError_t gemm(const GemmArgs& g, hipStream_t st) {
    const dim3 grid((g.N + kTN - 1) / kTN, (g.M + kTM - 1) / kTM);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, st, g);
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
__global__ void colsum_final_kernel(const float* part, int chunks, int N, float* out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int c = 0; c < chunks; c++) s += part[(long)c * N + n];
    out[n] = s;
}

__global__ void to_bf16_kernel(const float* x, short* y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// MSE: loss = mean (p - y)^2, dP = 2 (p - y) / n (FP32 and BF16)
__global__ void mse_kernel(const float* p, const float* y, long n, float* dp, short* dpb, float* part) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.f;
    if (i < n) {
        const float d = p[i] - y[i];
        e2 = d * d;
        const float gd = 2.f * d / (float)n;
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

// rows of a (n_rows x row_len) FP32 array picked by index, as BF16
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

// layer table: (rows per sample, in, out)
struct Layer {
    int rows, in, out;
};
constexpr Layer kL[7] = {{32, 512, 128}, {8, 512, 256}, {1, 2048, 512}, {1, 512, 4096},
                         {64, 64, 128},  {256, 32, 64}, {1024, 16, 4}};
constexpr int kG = 64, kT = 256, kMap = 64;
constexpr int kIn = kG * kT, kOut = kMap * kMap;

}  // namespace

struct cliffs_sur {
    int dev = 0, bmax = 0;
    hipStream_t st = nullptr;
    long np = 0;                         // parameter count
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 [np]
    short* wb = nullptr;                 // BF16 copy [np]
    long woff[7], boff[7];               // offsets of W_l (in x out, row-major) and b_l
    short* act[8] = {};                  // BF16 activations: act[0] input, act[l+1] output of L_l
    float* pred = nullptr;               // FP32 output (bmax * 4096)
    float* yb = nullptr;                 // targets of the batch
    float *dF = nullptr;                 // FP32 gradient scratch (largest activation)
    short *dA = nullptr, *dB = nullptr;  // BF16 gradient ping-pong
    float* part = nullptr;               // reduction partials
    float* hpart = nullptr;              // pinned
    int* didx = nullptr;
    float *dx = nullptr, *dy = nullptr;  // training data on the device
    long ndata = 0;
    int t = 0;                           // Adam step
};

namespace {

long act_elems(int l, int b) {  // elements of act[l] for batch b
    if (l == 0) return (long)b * kIn;
    return (long)b * kL[l - 1].rows * kL[l - 1].out;
}

int forward(cliffs_sur* s, int b) {
    for (int l = 0; l < 7; l++) {
        const Layer& L = kL[l];
        GemmArgs a{};
        a.M = b * L.rows;
        a.N = L.out;
        a.K = L.in;
        a.A = s->act[l];
        a.sam = L.in;
        a.sak = 1;
        a.B = s->wb + s->woff[l];
        a.sbk = L.out;
        a.sbn = 1;
        a.ldc = L.out;
        a.bias = s->w + s->boff[l];
        a.relu = l < 6;
        a.Cb = l < 6 ? s->act[l + 1] : nullptr;
        a.C = l < 6 ? nullptr : s->pred;
        CHK(gemm(a, s->st));
    }
    return CLIFFS_SUR_OK;
}

int colsum(cliffs_sur* s, const float* X, int M, int N, float* out) {
    const int chunks = (M + 255) / 256;
    hipLaunchKernelGGL(colsum_partial_kernel, dim3((N + 63) / 64, chunks), dim3(256), 0, s->st, X, M,
                       N, s->part);
    CHK(hipGetLastError());
    hipLaunchKernelGGL(colsum_final_kernel, dim3((N + 255) / 256), dim3(256), 0, s->st, s->part,
                       chunks, N, out);
    CHK(hipGetLastError());
    return CLIFFS_SUR_OK;
}

// backward from dP (FP32 in dF, BF16 in dA): parameter gradients into s->g
int backward(cliffs_sur* s, int b) {
    short* dcur = s->dA;   // BF16 gradient of layer l's output (pre-activation)
    short* dnext = s->dB;
    int rc;
    for (int l = 6; l >= 0; l--) {
        const Layer& L = kL[l];
        const int M = b * L.rows;
        // dZ_l is in dF (FP32) and dcur (BF16); bias gradient
        if ((rc = colsum(s, s->dF, M, L.out, s->g + s->boff[l]))) return rc;
        // dW_l = X_l^T dZ_l: A(m=i, k=r) = X[r*in + i]
        GemmArgs gw{};
        gw.M = L.in;
        gw.N = L.out;
        gw.K = M;
        gw.A = s->act[l];
        gw.sam = 1;
        gw.sak = L.in;
        gw.B = dcur;
        gw.sbk = L.out;
        gw.sbn = 1;
        gw.C = s->g + s->woff[l];
        gw.ldc = L.out;
        CHK(gemm(gw, s->st));
        if (l == 0) break;
        // dX_l = dZ_l W_l^T, masked by ReLU'(previous activation) = dZ_{l-1}
        GemmArgs gx{};
        gx.M = M;
        gx.N = L.in;
        gx.K = L.out;
        gx.A = dcur;
        gx.sam = L.out;
        gx.sak = 1;
        gx.B = s->wb + s->woff[l];
        gx.sbk = 1;
        gx.sbn = L.out;
        gx.C = s->dF;
        gx.Cb = dnext;
        gx.ldc = L.in;
        gx.mask = s->act[l];
        CHK(gemm(gx, s->st));
        short* tmp = dcur;
        dcur = dnext;
        dnext = tmp;
    }
    return CLIFFS_SUR_OK;
}

// loss and dP for the batch in act[0] / yb
int loss_grad(cliffs_sur* s, int b, float* loss) {
    const long n = (long)b * kOut;
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL(mse_kernel, dim3(blocks), dim3(256), 0, s->st, s->pred, s->yb, n, s->dF, s->dA,
                       s->part);
    CHK(hipGetLastError());
    CHK(hipMemcpyAsync(s->hpart, s->part, sizeof(float) * blocks, hipMemcpyDeviceToHost, s->st));
    CHK(hipStreamSynchronize(s->st));
    double sum = 0;
    for (int i = 0; i < blocks; i++) sum += s->hpart[i];
    *loss = (float)(sum / (double)n);
    return CLIFFS_SUR_OK;
}

int upload_batch(cliffs_sur* s, int b, const float* x, const float* y) {
    std::vector<float> tmp;
    CHK(hipMemcpyAsync(s->dF, x, sizeof(float) * (size_t)b * kIn, hipMemcpyHostToDevice, s->st));
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)(((long)b * kIn + 255) / 256)), dim3(256), 0,
                       s->st, s->dF, s->act[0], (long)b * kIn);
    CHK(hipGetLastError());
    if (y) CHK(hipMemcpyAsync(s->yb, y, sizeof(float) * (size_t)b * kOut, hipMemcpyHostToDevice, s->st));
    return CLIFFS_SUR_OK;
}

void free_sur(cliffs_sur* s) {
    if (!s) return;
    (void)hipSetDevice(s->dev);
    if (s->st) (void)hipStreamSynchronize(s->st);
    for (void* p : {(void*)s->w, (void*)s->g, (void*)s->m, (void*)s->v, (void*)s->wb, (void*)s->pred,
                    (void*)s->yb, (void*)s->dF, (void*)s->dA, (void*)s->dB, (void*)s->part,
                    (void*)s->didx, (void*)s->dx, (void*)s->dy})
        (void)hipFree(p);
    for (auto p : s->act) (void)hipFree(p);
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

}  // namespace

extern "C" {

const char* cliffs_sur_last_error(void) { return g_err.c_str(); }

int cliffs_sur_dims(int* gauges, int* samples, int* map, int* layers) {
    if (gauges) *gauges = kG;
    if (samples) *samples = kT;
    if (map) *map = kMap;
    if (layers) *layers = 7;
    return CLIFFS_SUR_OK;
}

int cliffs_sur_create(int device, int batch_max, unsigned seed, cliffs_sur** out) {
    if (!out || batch_max < 1) return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
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
    for (int l = 0; l < 7; l++) {
        s->woff[l] = np;
        np += (long)kL[l].in * kL[l].out;
        s->boff[l] = np;
        np += kL[l].out;
    }
    s->np = np;
    long amax = 0;
    for (int l = 0; l <= 7; l++) amax = std::max(amax, act_elems(l, batch_max));
    bool ok = hipMalloc(&s->w, sizeof(float) * np) == hipSuccess &&
              hipMalloc(&s->g, sizeof(float) * np) == hipSuccess &&
              hipMalloc(&s->m, sizeof(float) * np) == hipSuccess &&
              hipMalloc(&s->v, sizeof(float) * np) == hipSuccess &&
              hipMalloc(&s->wb, sizeof(short) * np) == hipSuccess &&
              hipMalloc(&s->pred, sizeof(float) * (size_t)batch_max * kOut) == hipSuccess &&
              hipMalloc(&s->yb, sizeof(float) * (size_t)batch_max * kOut) == hipSuccess &&
              hipMalloc(&s->dF, sizeof(float) * amax) == hipSuccess &&
              hipMalloc(&s->dA, sizeof(short) * amax) == hipSuccess &&
              hipMalloc(&s->dB, sizeof(short) * amax) == hipSuccess &&
              hipMalloc(&s->part, sizeof(float) * (amax / 64 + 4096)) == hipSuccess &&
              hipHostMalloc(&s->hpart, sizeof(float) * (amax / 64 + 4096)) == hipSuccess &&
              hipMalloc(&s->didx, sizeof(int) * batch_max) == hipSuccess;
    for (int l = 0; l <= 7 && ok; l++)
        ok = hipMalloc(&s->act[l], sizeof(short) * act_elems(l, batch_max)) == hipSuccess;
    if (!ok) return bail(fail(CLIFFS_SUR_ERR_NOMEM, "device allocation"));
    // He-uniform weights, zero biases (host PRNG: splitmix64)
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
    for (int l = 0; l < 7; l++) {
        const double lim = std::sqrt(6.0 / kL[l].in) * (l == 6 ? 0.1 : 1.0);
        for (long i = 0; i < (long)kL[l].in * kL[l].out; i++)
            hw[s->woff[l] + i] = (float)((2.0 * rnd() - 1.0) * lim);
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

int cliffs_sur_forward(cliffs_sur* s, int b, const float* x, float* y) {
    if (!s || !x || !y || b < 1 || b > s->bmax) return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
    CHK(hipSetDevice(s->dev));
    int rc;
    if ((rc = upload_batch(s, b, x, nullptr)) || (rc = forward(s, b))) return rc;
    CHK(hipMemcpyAsync(y, s->pred, sizeof(float) * (size_t)b * kOut, hipMemcpyDeviceToHost, s->st));
    CHK(hipStreamSynchronize(s->st));
    return CLIFFS_SUR_OK;
}

int cliffs_sur_loss_grad(cliffs_sur* s, int b, const float* x, const float* y, float* loss,
                         float* grads) {
    if (!s || !x || !y || !loss || b < 1 || b > s->bmax) return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
    CHK(hipSetDevice(s->dev));
    int rc;
    if ((rc = upload_batch(s, b, x, y)) || (rc = forward(s, b)) || (rc = loss_grad(s, b, loss)) ||
        (rc = backward(s, b)))
        return rc;
    if (grads) CHK(hipMemcpy(grads, s->g, sizeof(float) * s->np, hipMemcpyDeviceToHost));
    CHK(hipStreamSynchronize(s->st));
    return CLIFFS_SUR_OK;
}

int cliffs_sur_train(cliffs_sur* s, long n, const float* x, const float* y, int b, int steps, float lr,
                     unsigned seed, float* losses) {
    if (!s || n < 1 || !x || !y || b < 1 || b > s->bmax || steps < 0)
        return fail(CLIFFS_SUR_ERR_ARG, "bad arguments");
    CHK(hipSetDevice(s->dev));
    if (n != s->ndata) {
        (void)hipFree(s->dx);
        (void)hipFree(s->dy);
        s->dx = s->dy = nullptr;
        if (hipMalloc(&s->dx, sizeof(float) * n * kIn) != hipSuccess ||
            hipMalloc(&s->dy, sizeof(float) * n * kOut) != hipSuccess)
            return fail(CLIFFS_SUR_ERR_NOMEM, "training data allocation");
        s->ndata = n;
    }
    CHK(hipMemcpy(s->dx, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice));
    CHK(hipMemcpy(s->dy, y, sizeof(float) * n * kOut, hipMemcpyHostToDevice));
    unsigned long long z = 0xD1B54A32D192ED03ull * (seed + 7);
    std::vector<int> perm(n), idx(b);
    for (long i = 0; i < n; i++) perm[i] = (int)i;
    long pos = n;
    int rc;
    for (int it = 0; it < steps; it++) {
        for (int k = 0; k < b; k++) {  // epochs of a shuffled order (Fisher-Yates)
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
        CHK(hipMemcpyAsync(s->didx, idx.data(), sizeof(int) * b, hipMemcpyHostToDevice, s->st));
        hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)(((long)b * kIn + 255) / 256)), dim3(256), 0,
                           s->st, s->dx, s->didx, b, (long)kIn, s->act[0], (float*)nullptr);
        CHK(hipGetLastError());
        hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)(((long)b * kOut + 255) / 256)), dim3(256), 0,
                           s->st, s->dy, s->didx, b, (long)kOut, (short*)nullptr, s->yb);
        CHK(hipGetLastError());
        float loss = 0.f;
        if ((rc = forward(s, b)) || (rc = loss_grad(s, b, &loss)) || (rc = backward(s, b)) ||
            (rc = adam_all(s, lr)))
            return rc;
        if (losses) losses[it] = loss;
    }
    CHK(hipStreamSynchronize(s->st));
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
    if (hipMalloc(&fa, sizeof(float) * na) != hipSuccess || hipMalloc(&fb, sizeof(float) * nb) != hipSuccess ||
        hipMalloc(&fc, sizeof(float) * (size_t)M * N) != hipSuccess ||
        hipMalloc(&ba, sizeof(short) * na) != hipSuccess || hipMalloc(&bb, sizeof(short) * nb) != hipSuccess) {
        cleanup();
        return fail(CLIFFS_SUR_ERR_NOMEM, "device allocation");
    }
    hipError_t e = hipMemcpy(fa, A, sizeof(float) * na, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(fb, B, sizeof(float) * nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, nullptr, fa, ba, na);
        hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, nullptr, fb, bb, nb);
        GemmArgs g{};
        g.M = M;
        g.N = N;
        g.K = K;
        g.A = ba;
        g.sam = sam;
        g.sak = sak;
        g.B = bb;
        g.sbk = sbk;
        g.sbn = sbn;
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
