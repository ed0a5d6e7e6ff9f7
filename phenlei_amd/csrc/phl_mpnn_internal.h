// phl_mpnn_internal.h -- launchers of the warm-start network kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace phl {
namespace nn {

constexpr int kIn = 8;        // cell features
constexpr int kHidden = 128;  // hidden width
constexpr int kOut = 5;       // predicted state

// out[N,128] = A1[N,K1] W1 (+ A2[N,128] W2) + bias, epilogue by mode (see the
// kernel); transW takes W as [128,K] (the backward products dZ W^T).
void dense(const float *A1, const float *W1, int K1, const float *A2, const float *W2, const float *bias, int N,
           int mode, const float *Hres, float *Z, float *out, bool transW, hipStream_t s);
// The BF16 inference form of a hidden layer: out = h + relu(h A + m B + c)
// with BF16 products (FP32 accumulation); At, Bt from weights_bf16.
void weights_bf16(const float *A, const float *B, int L, __bf16 *out, hipStream_t s);
void dense_bf16(const float *h, const __bf16 *At, const float *m, const __bf16 *Bt, const float *bias, int N,
                float *out, hipStream_t s);
int grad_parts(int N);
// out[128,128] = A[N,128]^T D[N,128] (part: grad_parts(N) * 128 * 128 floats)
void wgrad(const float *A, const float *D, int N, float *part, float *out, hipStream_t s);
// out[K,C] = A[N,K]^T D[N,C] (A null: a column of ones, K = 1)
void small_grad(const float *A, int K, const float *D, int C, int N, float *part, float *out, hipStream_t s);
void neighbour_mean(const int *nbr, int W, int N, const float *invDeg, const float *h, float *m, hipStream_t s);
void neighbour_mean_transpose_add(const int *nbr, int W, int N, const float *invDeg, const float *dm, float *dh,
                                  hipStream_t s);
void relu_back(const float *dh, const float *z, size_t n, float *dz, hipStream_t s);
void add(float *a, const float *b, size_t n, hipStream_t s);
void out_layer(const float *h, const float *Wo, const float *bo, int N, float *y, hipStream_t s);
void out_layer_back(const float *dy, const float *Wo, int N, float *dh, hipStream_t s);
int loss_parts(int n);
void loss(const float *y, const float *t, int n, float *dy, double *part, double *out, hipStream_t s);
void adam(float *p, const float *g, float *m, float *v, int n, float lr, int step, hipStream_t s);

}  // namespace nn
}  // namespace phl
