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
 * M and N multiples of 64, K a multiple of 32.  BF16 products with K a
 * multiple of 64, 16-byte aligned operands and leading dimensions a multiple
 * of 8 run on the 16 x 16 x 32 MFMA kernel with the tile enc_gemm_tile
 * reports; the others on the generic 64 x 64 kernel.  An ENC_EPI_STORE
 * product on the MFMA kernel with fewer than 512 tiles splits K over up to 8
 * workgroups per tile: partial products go to a 16 MiB workspace of the
 * library (one per device) and are summed in split order, so the result does
 * not depend on scheduling, but two such calls on one device must not run at
 * the same time (different streams). */
int enc_gemm(int dtype, int ta, int tb, int epi, const void* A, const void* B, int M, int N, int K, int lda, int ldb,
             float* cf, void* ct, const float* bias, const float* zin, int ldc, void* stream);

/* The C tile (*bm x *bn rows x columns) enc_gemm's MFMA kernel takes for an
 * M x N x K product of dtype with epilogue epi (aligned operands): the largest
 * of 128 x 128, 128 x 64 and 64 x 128 that divides M and N in at least 512
 * tiles, else 64 x 64; 0 x 0 when the generic kernel runs (FP32, or K not a
 * multiple of 64).  *splits: the K splits (1, or for ENC_EPI_STORE under 512
 * tiles the largest power of two <= 8 keeping >= 4 stages of 64 per split, at
 * most 1,024 workgroups and the slabs within the workspace).  No device work. */
int enc_gemm_tile(int dtype, int epi, int M, int N, int K, int* bm, int* bn, int* splits);

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

/* The 2-layer network y = W2 gelu(W1 x + b1) + b2 (widths multiples of 64)
 * with its FP32 masters, the dtype copies the GEMMs read and Adam moments. */
typedef struct enc_mlp {
    int32_t dtype, d_in, hid, d_out;
    float *W1, *b1, *W2, *b2;          /* masters: d_in x hid, hid, hid x d_out, d_out */
    void *W1c, *W2c;                   /* dtype copies of W1 and W2 */
    float *m[4], *v[4];                /* Adam moments of W1, b1, W2, b2 */
} enc_mlp;

/* Per-step buffers for a batch of B rows: Z, dyf, dzf, Y FP32 (B x hid, B x
 * d_out, B x hid, B x d_out); H, dyt, dzt in dtype; gradients FP32 shaped
 * like the masters. */
typedef struct enc_mlp_work {
    float *Z, *Y, *dyf, *dzf;
    void *H, *dyt, *dzt;
    float *gW1, *gb1, *gW2, *gb2;
} enc_mlp_work;

/* One training step on B rows (X: B x d_in in dtype, T: B x d_out FP32
 * targets, colw: d_out column weights): forward, loss and gradient
 * (enc_loss_grad: loss += the batch's weighted mean squared error), the four
 * gradients (enc_gemm, enc_colsum) and the Adam step t (enc_adam) of every
 * parameter, in that order on one stream: the launches of the separate entry
 * points in one call. */
int enc_train_step(const enc_mlp* net, const enc_mlp_work* w, const void* X, const float* T, const float* colw, int B,
                   float lr, float beta1, float beta2, float eps, int32_t t, double* loss, void* stream);

/* y = x in dtype (round to nearest even for BF16). */
int enc_cast(int dtype, const float* x, void* y, int64_t n, void* stream);

/* Columns of the per-user scalar block of enc_user_features. */
#define ENC_SCALAR_COLS 64

/* Per-item day statistics of n ratings (items[k] < n_items, days[k]):
 * day_mean[j] and day_inv_std[j] = 1 / max(std, 1) of item j's rating days
 * (population std, from exact 64-bit integer sums); both 0 for an item
 * without ratings.  work: 3 * n_items uint64 of scratch. */
int enc_item_day_stats(const uint32_t* items, const int16_t* days, int64_t n, int n_items, uint64_t* work,
                       float* day_mean, float* day_inv_std, void* stream);

/* Token features and set pooling per user u < U.  User u's training ratings
 * are k = off[u] .. off[u+1]-1 (item items[k], rating class ratings[k] in
 * 1..5, training target targets[k], day days[k]).  Each rating is a token of
 *   e_k = q[items[k]] (BF16 bits, row pitch ldq; elements >= f zero),
 *   o_k = one-hot of the rating class (all zero outside 1..5),
 *   d_k = (days[k] - day_mean[j]) * day_inv_std[j]   (per-item normalised day),
 *   w_k = targets[k] - mu - ibias[j]                   (residual vs the biases).
 * Over S = up to max_items ratings at positions off[u] + floor(t cnt / |S|),
 * t < |S| (cnt the user's count), row u of out (pitch ldo >= 7 f + 64) gets
 *   [b f, b f + f)  mean_S o_k[b] e_k   for b = 0..4 (rating classes 1..5)
 *   [5 f, 6 f)      mean_S d_k e_k
 *   [6 f, 7 f)      mean_S w_k e_k
 * and, over all cnt ratings, the scalars at 7 f + c:
 *   c = 0 log1p(cnt) / 8;  1..5 share of rating class c;  6, 7 mean and std
 *   of w;  8, 9 mean and std of d;  10 mean day / 2243;  11 (max - min day) /
 *   2243;  12 min day / 2243;  13 max day / 2243;  14 the constant 1;
 *   15..63 zero.
 * A user without ratings gets zeros and the constant.  f <= 1024; ldq a
 * multiple of V and >= 64 V, V the smallest power of two with 64 V >= f. */
int enc_user_features(const uint16_t* q, int f, int ldq, const float* ibias, float mu, const float* day_mean,
                      const float* day_inv_std, const uint32_t* off, const uint32_t* items, const int8_t* ratings,
                      const float* targets, const int16_t* days, int U, int max_items, float* out, int64_t ldo,
                      void* stream);

/* out[r * ldo + c] = (x[rows[r] * ldx + c] - shift[c]) * scale[c] in dtype for
 * r < n, c < d (rows NULL: rows[r] = r; shift / scale NULL: 0 / 1), and 0 in
 * the padding n <= r < n_out or d <= c < ldo. */
int enc_affine_rows(int dtype, const float* x, int64_t ldx, const int64_t* rows, int n, int d, const float* shift,
                    const float* scale, void* out, int64_t ldo, int n_out, void* stream);

/* Column mean and 1 / std (population std; 1 where std <= 1e-6) of x (M x N,
 * pitch ld), with FP64 sums.  work: 512 * N doubles of scratch. */
int enc_colstats(const float* x, int M, int N, int64_t ld, double* work, float* mean, float* inv_std, void* stream);

const char* enc_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TSVD_ENC_H */
