/* intc_mlp.h -- C-ABI of the quartet-magnitude network in libintc.so.
 *
 * No counterpart in the reference: the reference screens with Schwarz
 * bounds only.  The network (12 -> 64 -> 64 -> 1, ReLU) predicts
 * log10(max|integral| / (Q_bra Q_ket)) of a contracted quartet from an
 * 8-float-per-pair feature table, so that the batched evaluator can skip
 * quartets below a calibrated threshold and route quartets between the FP64
 * and FP32-recurrence kernels (route codes of intc_plan_execute_routed).
 *
 * All pointers are device pointers; stream is a hipStream_t or NULL.
 * Parameters: NPARAM FP32 master weights (layout W1[64][16], b1[64],
 * W2[64][64], b2[64], w3[64], b3) and their BF16 copy (W1, W2, w3).
 * Return value: 0 on success, -1 on bad arguments, -3 on a launch failure.
 */
#ifndef INTC_MLP_H
#define INTC_MLP_H

#ifdef __cplusplus
extern "C" {
#endif

int intc_mlp_num_params(void);  /* FP32 master parameters */
int intc_mlp_num_bf16(void);    /* BF16 weights */
int intc_mlp_train_block(void); /* samples per training block */

/* X[n][16]: standardised features of quartets (quartet_pairs[2q] | quartet_pairs[2q+1]),
 * norm = 12 means then 12 scales; columns 12..15 are zero. */
int intc_mlp_features(long long n, const float *pair_feat, const int *quartet_pairs, const float *norm,
                      float *X, void *stream);

/* Inference and routing: pred[q] = log10 Q_bra Q_ket + min(0, network(q)) (pred may be NULL);
 * route[q] = 2 if pred + m_skip < log_skip, else 1 if pred + m_32 < log_32 and
 * log10 Q_bra Q_ket < log_qq32, else 0 (route may be NULL).  pair_feat must be
 * 16-byte and quartet_pairs 8-byte aligned. */
int intc_mlp_route(long long n, const float *pair_feat, const int *quartet_pairs, const float *norm,
                   const unsigned short *wb, const float *master, float m_skip, float log_skip, float m_32,
                   float log_32, float log_qq32, float *pred, signed char *route, void *stream);

/* One minibatch (samples idx[0..nb) of X/T): per-block partial gradients
 * (ceil(nb / intc_mlp_train_block()) x (NPARAM + 1) floats) and the weighted
 * squared-error loss added to *loss (under-predictions weigh under_weight). */
int intc_mlp_train_step(int nb, const float *X, const float *T, const int *idx, const unsigned short *wb,
                        const float *master, float *partial, float *loss, float under_weight, void *stream);

/* Sums nblk partial gradients (into grad_out when it is not NULL), applies
 * Adam to the master weights unless lr <= 0, and refreshes the BF16 copy. */
int intc_mlp_adam(int nblk, const float *partial, float *master, float *m1, float *m2, unsigned short *wb,
                  float lr, float beta1, float beta2, float eps, int step, float *grad_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
