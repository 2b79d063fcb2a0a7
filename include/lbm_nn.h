/*
 * lbm_nn.h -- C ABI of the kernels behind the learned warm start.
 *
 * The reference initialises a run from uniform fields or from a macro file
 * (S/Iterate.cu:211-275, S/FilesReading.cu:224-268) and has no surrogate:
 * these entry points are new.  They train and run a 3x3-convolution
 * encoder-decoder whose prediction (rho, u, v) is injected through
 * lbm_init / lbm_batch_init as f = f_eq(rho, u).
 *
 * All pointers are device pointers; `stream` is a hipStream_t (NULL: the
 * default stream).  Activations are FP32 NHWC; GEMM operands are bf16 (raw
 * 16-bit words); accumulation is FP32.  Calls enqueue work and return 0, or
 * 3 for bad arguments and 4 for a launch error.
 */
#ifndef LBM_NN_H
#define LBM_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NN_RELU 0x1   /* epilogue max(0, .) */
#define NN_ATOMIC 0x2 /* add the result into C (FP32 atomics); required with splits > 1 */
#define NN_TRANSC 0x4 /* store C transposed: element (m, n) at C[n*ldc + m] */

/* C(m,n) (+)= sum_r A(m,r) B(n,r) (+ bias[n]); A(m,r) = A[m*lda + r], or
 * A[r*lda + m] with ta; B(n,r) = B[n*ldb + r], or B[r*ldb + n] with tb.
 * splits > 1 divides the reduction over workgroups (needs NN_ATOMIC). */
int nn_gemm_bf16(int M, int N, int R, const void* A, int lda, int ta, const void* B, int ldb, int tb, float* C,
                 int ldc, const float* bias, int flags, int splits, void* stream);
/* im2col of a 3x3 same-padded convolution: col[p][(ky*3+kx)*C + c] (bf16),
 * zero for columns 9C..Kp-1 */
int nn_im2col3x3(const float* x, int Bn, int H, int W, int C, int Kp, void* col, void* stream);
/* adjoint of nn_im2col3x3 for an FP32 dcol: dx = col^T(dcol) */
int nn_col2im3x3(const float* dcol, int Bn, int H, int W, int C, int Kp, float* dx, void* stream);
/* dz = dy (times y > 0 when relu) as bf16; db[n] += column sums of the same */
int nn_grad_bf16(const float* dy, const float* y, int64_t n, int relu, void* dz, void* stream);
int nn_colsum(const float* dy, const float* y, int relu, int M, int N, float* db, void* stream);
int nn_to_bf16(const float* x, int64_t n, void* y, void* stream);
/* 2x2 average pooling and nearest x2 upsampling with their adjoints */
int nn_avgpool2(const float* x, int Bn, int H, int W, int C, float* y, void* stream);
int nn_avgpool2_bwd(const float* dy, int Bn, int H, int W, int C, float* dx, void* stream);
int nn_up2(const float* x, int Bn, int H, int W, int C, float* y, void* stream);
int nn_up2_bwd(const float* dy, int Bn, int H, int W, int C, float* dx, void* stream);
/* y = x + (1/W) sum_{x' >= x} x along the width (bwd = 0), or its adjoint
 * dx = dy + (1/W) sum_{x' <= x} dy (bwd = 1) */
int nn_xint(const float* x, int Bn, int H, int W, int C, int bwd, float* y, void* stream);
/* y = x + (1/H) sum over the height of x (symmetric: also its own adjoint) */
int nn_ymix(const float* x, int Bn, int H, int W, int C, float* y, void* stream);
int nn_axpy(float a, const float* x, int64_t n, float* y, void* stream);
int nn_add(const float* x, const float* y, int64_t n, float* z, void* stream);   /* z = x + y */
int nn_fill(float v, int64_t n, float* y, void* stream);
/* masked, channel-weighted MSE: *loss += sum w_c (pred - target)^2 * inv_count over
 * pixels with mask > 0; grad = its derivative */
int nn_mse(const float* pred, const float* target, const float* mask, const float* wch, int64_t npix, int C,
           float inv_count, float* grad, float* loss, void* stream);
/* Adam step t (>= 1) on FP32 weights w with moments m, v; the first n_bf
 * entries are also written to wbf as bf16 */
int nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
            int64_t t, void* wbf, int64_t n_bf, void* stream);

#ifdef __cplusplus
}
#endif
#endif
