// Host FP64 dense linear algebra used by the SCF driver in place of the reference's OpenBLAS/LAPACK
// calls (src/linear_algebra.zig: mmSlice -> cblas_dgemm :122-147, geighSlice -> LAPACKE_dsygvd :184-200,
// eighSlice -> dsyevd :55-70, luFactorize/luSolve -> dgetrf/dgetrs :203-231).  Row-major throughout.
#pragma once

#include <cstddef>

namespace zq {

// C = alpha * op(A) op(B) + beta * C for square n x n row-major matrices (op = identity).
void gemm(int n, const double *A, const double *B, double *C, double alpha = 1.0, double beta = 0.0);

// Symmetric eigenproblem: A (n x n) is replaced by eigenvectors in its columns, w by eigenvalues in
// ascending order.  Householder tridiagonalisation followed by implicit QL with Wilkinson shifts.
// Returns 0, or 1 if the QL iteration failed to converge.
int syev(int n, double *A, double *w);

// Generalised symmetric-definite problem A x = w B x (itype 1): Cholesky of B, reduction, syev,
// back-transformation; eigenvectors in the columns of A, normalised to x^T B x = 1.  B is replaced by
// its Cholesky factor.  Returns 0, 1 (QL failure) or 2 (B not positive definite).
int sygv(int n, double *A, double *B, double *w);

// X = L^-T for the Cholesky factor S = L L^T (so X^T S X = I); 0, or 1 if S is not positive definite.
int chol_inv_t(int n, const double *S, double *X);

// LU with partial pivoting; 0 on success, 1 on an exactly singular pivot (dgetrf info > 0).
int getrf(int n, double *A, int *piv);
void getrs(int n, const double *LU, const int *piv, double *b);

}  // namespace zq
