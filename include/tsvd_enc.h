/* Fold-in encoder kernels: set-pooled item embeddings -> 2-layer MLP (GELU)
 * -> initial user factors p_u and user bias b_u for users folded into a
 * trained tsvdx4 model.
 *
 * The reference has no counterpart: it starts every user from random factors
 * (TxModel::new, src/tx.rs:460-649) or zeroes them at reset_u_epoch
 * (src/tx.rs:745-748); its closest analogue of fold-in is the user-only pass
 * of Tsvdx4pModel (fit_epoch_user, src/tsvdx4p.rs:390-639), exposed by
 * tx_fit_users_epoch in tsvd_tx.h.  This library supplies the warm start for
 * that pass.
 *
 * Every pointer is device memory of the calling process (any allocator), and
 * `stream` is a hipStream_t (NULL: the default stream).  Matrices are
 * row-major.  dtype: ENC_BF16 (bfloat16 bits in uint16, MFMA BF16 with FP32
 * accumulation) or ENC_F32 (FP32 MFMA, the debug build used for tight
 * gradient checks).  Functions return 0 or a negative code; enc_last_error()
 * describes the failure (thread-local). */
#ifndef TSVD_ENC_H
#define TSVD_ENC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum enc_dtype { ENC_BF16 = 0, ENC_F32 = 1 };

/* Epilogues of enc_gemm (acc = the FP32 product of row r, column c):
 *   ENC_EPI_STORE:     cf[r*ldc + c] = acc
 *   ENC_EPI_BIAS:      cf[r*ldc + c] = acc + bias[c]
 *   ENC_EPI_BIAS_GELU: z = acc + bias[c]; cf[...] = z; ct[...] = gelu(z)
 *   ENC_EPI_GELU_BWD:  g = acc * gelu'(zin[r*ldc + c]); cf[...] = g; ct[...] = g
 * gelu(z) = z * Phi(z) (exact, erf form); ct has element type dtype. */
enum enc_epi { ENC_EPI_STORE = 0, ENC_EPI_BIAS = 1, ENC_EPI_BIAS_GELU = 2, ENC_EPI_GELU_BWD = 3 };

/* C (M x N) = op(A) (M x K) * op(B) (K x N) with FP32 accumulation, where
 * op(A) = A (A stored M x K, leading dimension lda) or, with ta, A^T (A stored
 * K x M); likewise op(B) = B (K x N, ldb) or, with tb, B^T (B stored N x K).
 * M and N multiples of 64, K a multiple of 32. */
int enc_gemm(int dtype, int ta, int tb, int epi, const void* A, const void* B, int M, int N, int K, int lda, int ldb,
             float* cf, void* ct, const float* bias, const float* zin, int ldc, void* stream);

/* Weighted squared error of Y (M x N, ld) against T: loss += sum_{r,c}
 * colw[c] (Y - T)^2 / M (FP64 accumulation into *loss); gradient dY = 2 colw[c]
 * (Y - T) / M written as FP32 (dyf) and as dtype (dyt). */
int enc_loss_grad(int dtype, const float* Y, const float* T, const float* colw, float* dyf, void* dyt, int M, int N,
                  int ld, double* loss, void* stream);

/* out[c] = sum_r X[r*ld + c] for c < N. */
int enc_colsum(const float* X, int M, int N, int ld, float* out, void* stream);

/* One Adam step on n FP32 master weights w with gradient g and moments m, v
 * (bias-corrected with step t >= 1); the updated weights are also written to
 * wcopy in dtype (the copy the GEMMs read). */
int enc_adam(int dtype, float* w, const float* g, float* m, float* v, void* wcopy, int64_t n, float lr, float beta1,
             float beta2, float eps, int32_t t, void* stream);

/* y = x in dtype (round to nearest even for BF16). */
int enc_cast(int dtype, const float* x, void* y, int64_t n, void* stream);

/* Set pooling per user u < U over its rated items items[off[u] .. off[u+1])
 * (up to max_items of them, taken at evenly spaced positions):
 *   out[u*ldo + k]      = mean_j q[j*fs + k]
 *   out[u*ldo + fs + k] = mean_j w_j q[j*fs + k]       (k < fs)
 * with w_j the weight of that rating; users without items get zeros. */
int enc_pool(const float* q, int fs, const uint32_t* off, const uint32_t* items, const float* w, int U, int max_items,
             float* out, int ldo, void* stream);

const char* enc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TSVD_ENC_H */
