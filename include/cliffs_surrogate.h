/*
 * cliffs_surrogate -- learned nest surrogate on MI355X (gfx950).
 *
 * A network that maps the surface-elevation series of 64 gauges on a nested
 * grid's boundary (512 samples each) to the grid's 512 x 512 map of maximum
 * inundation and to its maximum run-up, trained on runs the solver itself
 * generates.  The full nested solve is its fallback and the answer its
 * error is measured against.  No counterpart in the reference (Cliffs has
 * no learned component).
 *
 * Every layer is a GEMM on the MFMA units: BF16 operands, FP32 accumulation,
 * hand-written kernels for the forward pass, the backward pass and Adam.
 * Layers: four 1-D convolutions along time (E1: kernel 4 = stride 4, 64 ->
 * 128 channels; E2: kernel 5, stride 1, 128 -> 128; E3: kernel 5, stride 2,
 * 128 -> 256; E4: kernel 5, stride 2, 256 -> 256; zero padding of 2 rows
 * for kernel 5), maximum and mean over the remaining 32 time rows (512
 * features), dense D1, D2 (512 -> 512), the run-up head HR (512 -> 1), the
 * map head M1 (512 -> 4096 = 8 x 8 x 64) and six 2 x 2 stride-2 transposed
 * convolutions T1..T6 (64 -> 32 -> 16 -> 8 -> 8 -> 4 -> 1 channels); ReLU
 * after every layer but HR and T6.
 *
 * Layouts (FP32, host memory, row-major):
 *   x: (batch, 512 samples, 64 gauges)  -- time-major rows
 *   y: (batch, 262144) maps in quadtree order: index
 *        (((((p0 * 4 + s1) * 4 + s2) * 4 + s3) * 4 + s4) * 4 + s5) * 4 + s6
 *      with p0 = (iy >> 6) * 8 + (ix >> 6) the 8 x 8 base pixel and
 *      s_k = 2 * ((iy >> (6-k)) & 1) + ((ix >> (6-k)) & 1) the sub-pixel at
 *      level k, for pixel (iy, ix) of the 512 x 512 map.
 *   runup: (batch) one value per sample, in the units the caller trains it in.
 *   parameters: W, b of E1..E4, D1, D2, HR, M1, T1..T6 concatenated; W
 *      (in x out) row-major, in = kernel * channels for E1..E4 (window rows
 *      time-major: W[(dt * cin + c) * cout + o]).
 *
 * Every function returns CLIFFS_SUR_OK (0) or a negative error code;
 * cliffs_sur_last_error() describes the failure.
 */
#ifndef CLIFFS_SURROGATE_H
#define CLIFFS_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

#define CLIFFS_SUR_OK 0
#define CLIFFS_SUR_ERR_ARG (-1)
#define CLIFFS_SUR_ERR_HIP (-2)
#define CLIFFS_SUR_ERR_NOMEM (-3)
#define CLIFFS_SUR_ERR_NODEVICE (-5)

typedef struct cliffs_sur cliffs_sur;

const char *cliffs_sur_last_error(void);
/* 64 gauges, 512 samples, 512 x 512 map, 14 layers */
int cliffs_sur_dims(int *gauges, int *samples, int *map, int *layers);

/* Network with He-uniform weights from `seed`, for batches of up to
 * batch_max (<= 256) samples. */
int cliffs_sur_create(int device, int batch_max, unsigned seed, cliffs_sur **out);
int cliffs_sur_destroy(cliffs_sur *s);
int cliffs_sur_nparams(cliffs_sur *s, long *n);
int cliffs_sur_get_params(cliffs_sur *s, float *params);
/* Also resets the optimiser state. */
int cliffs_sur_set_params(cliffs_sur *s, const float *params);

/* Inference for `batch` samples: the map (y, or NULL) and the run-up
 * (runup, or NULL). */
int cliffs_sur_forward(cliffs_sur *s, int batch, const float *x, float *y, float *runup);
/* Loss w_map * mean squared map error + w_runup * mean squared run-up error
 * of net(x) against (y, yr) and, if grads is not NULL, its gradient with
 * respect to every parameter (same layout as the parameters). */
int cliffs_sur_loss_grad(cliffs_sur *s, int batch, const float *x, const float *y, const float *yr,
                         float w_map, float w_runup, float *loss, float *grads);
/* The training set (n samples of x, y, yr), copied to the device, where it
 * stays for cliffs_sur_train until replaced or the network is destroyed. */
int cliffs_sur_set_data(cliffs_sur *s, long n, const float *x, const float *y, const float *yr);
/* `steps` Adam steps (learning rate lr) on random batches of `batch` of the
 * training set, shuffled per epoch from `seed`; losses[k] = loss of step k
 * (or NULL). */
int cliffs_sur_train(cliffs_sur *s, int batch, int steps, float lr, float w_map, float w_runup,
                     unsigned seed, float *losses);

/* Ground heights of the map's cells (MAP x MAP FP32, raster order; any
 * value above every flood level, such as +inf, where a cell is not land),
 * for the inundation maps of cliffs_sur_predict. */
int cliffs_sur_set_ground(cliffs_sur *s, const float *ground);
/* The answer of k networks (on one device) for `batch` samples: each
 * network's run-up (runups[i * batch + j], or NULL) and, if maps is not
 * NULL, the inundation maps (batch x MAP x MAP, raster order) of the
 * networks' mean map output read as flood levels: the ground height (of
 * nets[0]'s ground) where ground <= level, 0 elsewhere. */
int cliffs_sur_predict(cliffs_sur *const *nets, int k, int batch, const float *x, float *runups,
                       float *maps);

/* The GEMM kernel alone: C (M x N, row-major FP32) = A B with the operands
 * rounded to BF16, A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]. */
int cliffs_sur_gemm_bf16(int device, int M, int N, int K, const float *A, long sam, long sak,
                         const float *B, long sbk, long sbn, float *C);

#ifdef __cplusplus
}
#endif

#endif /* CLIFFS_SURROGATE_H */
