/*
 * mgcm_sg.h -- C ABI of the learned IR heating-rate surrogate: an MLP
 * 147 -> 512 -> 512 -> 512 -> 36 (GELU) with BF16 weights and activations
 * and FP32 accumulation on the MI355X matrix cores, trained on the device by
 * its own forward and backward kernels.  The plain radiation path
 * (mgcm_rt.h) generates its training data and re-solves the columns whose
 * consistency check fails.
 *
 * No counterpart in the reference: it is a learned stand-in for the IR part of
 * SFLUXI/GFLUXI (code/sfluxi.f90, code/gfluxi.f90) as COMP3 uses it -- the IR
 * heating of each GCM layer (code/comp3.f:735-738,765-771).
 *
 * Features of a column (mgcm_sg_features), for GCM layers l = 1..36 (radiation
 * layer L = l+1, mid-level K = 2L): TMID(K), log10 PMID(K), log10 of the
 * layer's dust optical depth TAUREF(2L+1)+TAUREF(2L+2) (floored at 1e-12),
 * log10 QH2O(K) (floored at 1e-30); then TLEV(L_LEVELS)
 * (surface temperature), emissivity 1-ALBI and surface pressure PLEV(L_LEVELS).
 * Each feature is standardised with the mean/inverse-std arrays the caller
 * passes (statistics of the training set).
 *
 * Parameters: one FP32 vector of MGCM_SG_NPARAM values, per layer the weights
 * W[k][n] (inputs k, outputs n) followed by the bias b[n]: W1 is 160 x 512
 * (rows 147..159 are padding and stay zero), W2 and W3 512 x 512, W4 512 x 36.
 */
#ifndef MGCM_SG_H
#define MGCM_SG_H

#ifdef __cplusplus
extern "C" {
#endif

#define MGCM_SG_NLAY 36
#define MGCM_SG_NIN 147  /* 4 x 36 + 3 */
#define MGCM_SG_KIN 160  /* NIN padded to the matrix-core K step */
#define MGCM_SG_NHID 512
#define MGCM_SG_NOUT 36
#define MGCM_SG_NPARAM (160 * 512 + 512 + 2 * (512 * 512 + 512) + 512 * 36 + 36)

typedef struct mgcm_sg mgcm_sg;

/* Device buffers for batches of up to max_batch columns (max_batch a multiple
 * of 32 for training); parameters initialised from seed (He-normal weights,
 * zero biases), Adam state zero. */
int mgcm_sg_create(int max_batch, unsigned long long seed, mgcm_sg **out);
int mgcm_sg_destroy(mgcm_sg *sg);
int mgcm_sg_get_params(mgcm_sg *sg, float *params);       /* host, MGCM_SG_NPARAM */
int mgcm_sg_set_params(mgcm_sg *sg, const float *params); /* host; resets Adam */

/* x: device [batch][MGCM_SG_NIN] standardised features; y: device
 * [batch][MGCM_SG_NOUT].  Asynchronous on stream. */
int mgcm_sg_forward(mgcm_sg *sg, int batch, const float *x, float *y, void *stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the mean squared error
 * between the outputs and targets t (device [batch][MGCM_SG_NOUT]); batch a
 * multiple of 32.  loss (device, one float) receives the batch's loss. */
int mgcm_sg_train_step(mgcm_sg *sg, int batch, const float *x, const float *t, float lr, float *loss,
                       void *stream);

/* Gradients of that loss at the current parameters into grads (host,
 * MGCM_SG_NPARAM); parameters unchanged.  Synchronises. */
int mgcm_sg_gradients(mgcm_sg *sg, int batch, const float *x, const float *t, float *grads);

/* The consistency check of the surrogate form, on the device: for each of ncol
 * columns, the outputs of two networks y0, y1 ([ncol][MGCM_SG_NOUT]; y1 may be
 * NULL) de-standardised as h = scale * sinh(y * tstd + tmean) (FP64; tmean,
 * tstd [MGCM_SG_NOUT] FP64), h of y0 into h [ncol][MGCM_SG_NOUT], the largest
 * |h0 - h1| into dis [ncol], and into inr [ncol] 1 where every standardised
 * feature x [ncol][MGCM_SG_NIN] lies in [xlo, xhi] (FP32 [MGCM_SG_NIN]), else 0.
 * Asynchronous on stream. */
int mgcm_sg_check(int ncol, const float *y0, const float *y1, const float *x, const double *tmean,
                  const double *tstd, const float *xlo, const float *xhi, double scale, double *h, double *dis,
                  int *inr, void *stream);

/* Standardised features of ncol columns of a grid with nlay = 36 layers,
 * device pointers in the layouts of mgcm_rt_columns; mean/inv_std: device
 * [MGCM_SG_NIN]; x: device [ncol][MGCM_SG_NIN]. */
int mgcm_sg_features(int ncol, int nlay, const double *plev, const double *tlev, const double *pmid,
                     const double *tmid, const double *qh2o, const double *tauref, const double *albi,
                     const float *mean, const float *inv_std, float *x, void *stream);

/* Plain GEMM on the matrix cores, for checks: C[m][n] = sum_k A[m][k] B[k][n]
 * with A, B BF16 row-major, C FP32 row-major.  Device pointers; k a multiple
 * of 32.  Synchronises. */
int mgcm_sg_gemm_check(int m, int n, int k, const void *a, const void *b, float *c);

#ifdef __cplusplus
}
#endif
#endif
