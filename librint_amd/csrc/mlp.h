// Kernels of the learned SCF guess (mlp.hip).  All pointers are device
// pointers; BF16 values are raw 16-bit patterns (upper half of an FP32).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace librint {

// C[M x N] = act(op(A) op(B) + bias (+ C if accumulate)), act 0 none / 1 ReLU.
// op(A)(m, k) = transA ? A[k*lda + m] : A[m*lda + k]
// op(B)(k, n) = transB ? B[n*ldb + k] : B[k*ldb + n]
// C (FP32, may be null when Cb is given) and/or Cb (BF16 copy) are written.
hipError_t launch_gemm_bf16(int M, int N, int K, const uint16_t *A, int lda, int transA, const uint16_t *B, int ldb,
                            int transB, float *C, int ldc, const float *bias, int act, int accumulate, uint16_t *Cb,
                            int ldcb, hipStream_t st);
hipError_t launch_relu_bwd(size_t n, const float *dz, const float *h, float *dpre, uint16_t *dpre_b, hipStream_t st);
hipError_t launch_colsum(int rows, int cols, const float *x, int ld, float *out, int accumulate, hipStream_t st);
hipError_t launch_mse(int rows, int cols, const float *y, int ldy, const float *t, const float *mask, int ldt,
                      float scale, float *dy, uint16_t *dy_b, double *loss, hipStream_t st);
hipError_t launch_adam(size_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                       int step, uint16_t *wb, hipStream_t st);
hipError_t launch_f32_to_bf16(size_t n, const float *x, uint16_t *y, hipStream_t st);

}  // namespace librint
