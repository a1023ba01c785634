#pragma once

namespace librint {

// Symmetric eigen-decomposition A = V diag(w) V^T, w ascending,
// V[k*n + i] = component k of eigenvector i (row-major n x n).
void eigh(int n, const double *A, double *w, double *V);
// C = A B and C = A^T B for row-major n x n matrices.
void gemm_nn(int n, const double *A, const double *B, double *C);
void gemm_tn(int n, const double *A, const double *B, double *C);

}  // namespace librint
