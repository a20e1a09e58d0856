/* diskfft on MI355X -- surrogate MLP (no counterpart in the reference).
 *
 * An ensemble of `members` networks 4 -> 128 -> 128 -> 128 -> 3 with GELU
 * hidden layers, BF16 weights and activations and FP32 accumulation on the
 * matrix cores, trained with Adam by the library's own fused kernels.  Inputs
 * and targets are whatever the caller normalises them to (the surrogate maps
 * disk parameters to gamma, Omega_p and saturated |C2|/|C0|).
 *
 * Parameter vector per member (FP32, 34051 values, W[out][in] row-major):
 *   W1[128][4] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[3][128] b4[3]
 */
#ifndef DISKFFT_MLP_H
#define DISKFFT_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

#define DFK_MLP_OK 0
#define DFK_MLP_EINVAL 1
#define DFK_MLP_EHIP 2

typedef struct dfk_mlp dfk_mlp;

const char *dfk_mlp_last_error(void);
int dfk_mlp_param_count(void);

/* Glorot-uniform weights and zero biases from `seed` (member m uses its own
 * stream). */
int dfk_mlp_create(int members, int device, unsigned long long seed, dfk_mlp **out);
void dfk_mlp_destroy(dfk_mlp *h);
int dfk_mlp_get_params(dfk_mlp *h, int member, float *params);
int dfk_mlp_set_params(dfk_mlp *h, int member, const float *params);

/* y[member][i][3] for x[i][4], every member (fused inference kernel). */
int dfk_mlp_forward(dfk_mlp *h, const float *x, int n, float *y);

/* Loss (mean squared error over the n points and 3 outputs) and its gradient
 * with respect to every parameter, per member, from one fused forward +
 * backward pass over all n points; weights unchanged. */
int dfk_mlp_grad(dfk_mlp *h, const float *x, const float *t, int n, float *grad, float *loss);

/* Adam (beta 0.9 / 0.999, eps 1e-8) on minibatches of `batch` points,
 * reshuffled per member and epoch from `seed`.  loss_hist[epoch] (may be
 * NULL) receives the epoch's mean training loss averaged over members. */
int dfk_mlp_train(dfk_mlp *h, const float *x, const float *t, int n, int epochs, int batch,
                  float lr, unsigned long long seed, float *loss_hist);

/* D = A B for one 16 x 16 x 16 tile in BF16 with FP32 accumulation, through the
 * same matrix-core tile routine as the network kernels (row-major 16 x 16). */
int dfk_mlp_mfma_probe(const float *A, const float *B, float *D);

#ifdef __cplusplus
}
#endif
#endif
