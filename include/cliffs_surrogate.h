/*
 * cliffs_surrogate -- learned nest surrogate on MI355X (gfx950).
 *
 * A small network that maps the surface-elevation series of 64 gauges at a
 * nested grid's boundary (256 samples each) to the grid's 64 x 64 map of
 * maximum inundation, trained on runs the solver itself generates.  The full
 * nested solve is its fallback and the answer its error is measured
 * against.  No counterpart in the reference (Cliffs has no learned
 * component).
 *
 * Every layer is a GEMM on the MFMA units: BF16 operands, FP32 accumulation,
 * hand-written kernels for the forward pass, the backward pass and Adam.
 *
 * Layouts (FP32, host memory, row-major):
 *   x: (batch, 256 samples, 64 gauges)  -- time-major rows
 *   y: (batch, 4096) maps in quadtree order: index
 *        ((p0 * 4 + s1) * 4 + s2) * 4 + s3
 *      with p0 = (iy >> 3) * 8 + (ix >> 3) the 8 x 8 base pixel and
 *      s_k = 2 * ((iy >> (3-k)) & 1) + ((ix >> (3-k)) & 1) the sub-pixel at
 *      level k, for pixel (iy, ix) of the 64 x 64 map.
 *   parameters: W1, b1, ..., W7, b7 concatenated, W_l (in x out) row-major.
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
/* 64 gauges, 256 samples, 64 x 64 map, 7 layers */
int cliffs_sur_dims(int *gauges, int *samples, int *map, int *layers);

/* Network with He-uniform weights from `seed`, for batches up to batch_max. */
int cliffs_sur_create(int device, int batch_max, unsigned seed, cliffs_sur **out);
int cliffs_sur_destroy(cliffs_sur *s);
int cliffs_sur_nparams(cliffs_sur *s, long *n);
int cliffs_sur_get_params(cliffs_sur *s, float *params);
/* Also resets the optimiser state. */
int cliffs_sur_set_params(cliffs_sur *s, const float *params);

/* Inference: y = net(x) for `batch` samples. */
int cliffs_sur_forward(cliffs_sur *s, int batch, const float *x, float *y);
/* Mean squared error of net(x) against y and, if grads is not NULL, its
 * gradient with respect to every parameter (same layout as the parameters). */
int cliffs_sur_loss_grad(cliffs_sur *s, int batch, const float *x, const float *y, float *loss,
                         float *grads);
/* `steps` Adam steps (learning rate lr) on random batches of `batch` of the n
 * samples (x, y), shuffled per epoch from `seed`; losses[k] = loss of step k
 * (or NULL). */
int cliffs_sur_train(cliffs_sur *s, long n, const float *x, const float *y, int batch, int steps,
                     float lr, unsigned seed, float *losses);

/* The GEMM kernel alone: C (M x N, row-major FP32) = A B with the operands
 * rounded to BF16, A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]. */
int cliffs_sur_gemm_bf16(int device, int M, int N, int K, const float *A, long sam, long sak,
                         const float *B, long sbk, long sbn, float *C);

#ifdef __cplusplus
}
#endif

#endif /* CLIFFS_SURROGATE_H */
