// Host FP64 linear algebra for the SCF driver (see zq_linalg.h).  Algorithms: Householder reduction
// to tridiagonal form and the implicit QL method with Wilkinson shifts (Wilkinson & Reinsch, Handbook
// for Automatic Computation II, 1971; Golub & Van Loan, "Matrix Computations", 8.3), Cholesky
// reduction of the generalized problem (Golub & Van Loan 8.7.2).

#include "zq_linalg.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace zq {

void gemm(int n, const double *A, const double *B, double *C, double alpha, double beta) {
  std::vector<double> acc(size_t(n) * n, 0.0);
  constexpr int BS = 64;
  for (int i0 = 0; i0 < n; i0 += BS)
    for (int k0 = 0; k0 < n; k0 += BS)
      for (int i = i0; i < std::min(n, i0 + BS); ++i) {
        double *__restrict__ c = &acc[size_t(i) * n];
        for (int k = k0; k < std::min(n, k0 + BS); ++k) {
          const double a = A[size_t(i) * n + k];
          const double *__restrict__ b = &B[size_t(k) * n];
          for (int j = 0; j < n; ++j) c[j] += a * b[j];
        }
      }
  for (size_t x = 0; x < size_t(n) * n; ++x) C[x] = alpha * acc[x] + (beta == 0.0 ? 0.0 : beta * C[x]);
}

// Reduce symmetric A to tridiagonal T = Q^T A Q; on return A holds Q (row-major), d the diagonal and
// e[i] the coupling between i and i+1 (e[n-1] = 0).
static void tridiagonalize(int n, double *A, double *d, double *e) {
  std::vector<double> Q(size_t(n) * n, 0.0), v(n), p(n);
  for (int i = 0; i < n; ++i) Q[size_t(i) * n + i] = 1.0;
  for (int k = 0; k + 2 < n; ++k) {
    // Householder vector for column k below the diagonal
    double norm = 0;
    for (int i = k + 1; i < n; ++i) norm += A[size_t(i) * n + k] * A[size_t(i) * n + k];
    norm = std::sqrt(norm);
    if (norm == 0.0) continue;
    const double x0 = A[size_t(k + 1) * n + k];
    const double alpha = x0 > 0 ? -norm : norm;
    std::fill(v.begin(), v.end(), 0.0);
    v[k + 1] = x0 - alpha;
    for (int i = k + 2; i < n; ++i) v[i] = A[size_t(i) * n + k];
    double vn = 0;
    for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    const double beta = 2.0 / vn;
    // p = beta * A v (trailing block), K = beta/2 * v^T p, q = p - K v, A -= v q^T + q v^T
    for (int i = k; i < n; ++i) {
      double s = 0;
      const double *row = &A[size_t(i) * n];
      for (int j = k + 1; j < n; ++j) s += row[j] * v[j];
      p[i] = beta * s;
    }
    double K = 0;
    for (int i = k + 1; i < n; ++i) K += v[i] * p[i];
    K *= 0.5 * beta;
    for (int i = k + 1; i < n; ++i) p[i] -= K * v[i];
    // column/row k: A[k][k+1..] becomes alpha e1
    for (int i = k + 1; i < n; ++i) {
      double *row = &A[size_t(i) * n];
      for (int j = k + 1; j < n; ++j) row[j] -= v[i] * p[j] + p[i] * v[j];
    }
    A[size_t(k + 1) * n + k] = alpha;
    A[size_t(k) * n + k + 1] = alpha;
    for (int i = k + 2; i < n; ++i) A[size_t(i) * n + k] = A[size_t(k) * n + i] = 0.0;
    // Q <- Q (I - beta v v^T)
    for (int r = 0; r < n; ++r) {
      double *qr = &Q[size_t(r) * n];
      double s = 0;
      for (int j = k + 1; j < n; ++j) s += qr[j] * v[j];
      s *= beta;
      for (int j = k + 1; j < n; ++j) qr[j] -= s * v[j];
    }
  }
  for (int i = 0; i < n; ++i) {
    d[i] = A[size_t(i) * n + i];
    e[i] = (i + 1 < n) ? A[size_t(i + 1) * n + i] : 0.0;
  }
  std::memcpy(A, Q.data(), sizeof(double) * n * n);
}

// Implicit QL on the tridiagonal (d, e); rotations accumulated into the columns of Z (row-major).
static int tql(int n, double *d, double *e, double *Z) {
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= std::numeric_limits<double>::epsilon() * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) return 1;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      for (i = m - 1; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for (int k = 0; k < n; ++k) {
          double *zk = &Z[size_t(k) * n];
          f = zk[i + 1];
          zk[i + 1] = s * zk[i] + c * f;
          zk[i] = c * zk[i] - s * f;
        }
      }
      if (r == 0.0 && i >= l) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  return 0;
}

int syev(int n, double *A, double *w) {
  std::vector<double> e(n);
  tridiagonalize(n, A, w, e.data());
  if (tql(n, w, e.data(), A)) return 1;
  // sort ascending with the eigenvector columns
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return w[a] < w[b]; });
  std::vector<double> Z(A, A + size_t(n) * n), ws(w, w + n);
  for (int j = 0; j < n; ++j) {
    w[j] = ws[idx[j]];
    for (int k = 0; k < n; ++k) A[size_t(k) * n + j] = Z[size_t(k) * n + idx[j]];
  }
  return 0;
}

int sygv(int n, double *A, double *B, double *w) {
  // Cholesky B = L L^T (lower, in place)
  for (int j = 0; j < n; ++j) {
    double s = B[size_t(j) * n + j];
    for (int k = 0; k < j; ++k) s -= B[size_t(j) * n + k] * B[size_t(j) * n + k];
    if (s <= 0.0) return 2;
    const double ljj = std::sqrt(s);
    B[size_t(j) * n + j] = ljj;
    for (int i = j + 1; i < n; ++i) {
      double t = B[size_t(i) * n + j];
      for (int k = 0; k < j; ++k) t -= B[size_t(i) * n + k] * B[size_t(j) * n + k];
      B[size_t(i) * n + j] = t / ljj;
    }
  }
  const double *L = B;
  // W = L^-1 A (forward substitution on columns), A' = W L^-T
  std::vector<double> W(size_t(n) * n);
  for (int i = 0; i < n; ++i) {
    double *wi = &W[size_t(i) * n];
    std::memcpy(wi, &A[size_t(i) * n], sizeof(double) * n);
    for (int k = 0; k < i; ++k) {
      const double l = L[size_t(i) * n + k];
      const double *wk = &W[size_t(k) * n];
      for (int c = 0; c < n; ++c) wi[c] -= l * wk[c];
    }
    const double inv = 1.0 / L[size_t(i) * n + i];
    for (int c = 0; c < n; ++c) wi[c] *= inv;
  }
  // A'[r][i] = (W[r][i] - sum_{k<i} L[i][k] A'[r][k]) / L[i][i]
  for (int r = 0; r < n; ++r) {
    double *ar = &A[size_t(r) * n];
    const double *wr = &W[size_t(r) * n];
    for (int i = 0; i < n; ++i) {
      double t = wr[i];
      const double *li = &L[size_t(i) * n];
      for (int k = 0; k < i; ++k) t -= li[k] * ar[k];
      ar[i] = t / li[i];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) {
      const double avg = 0.5 * (A[size_t(i) * n + j] + A[size_t(j) * n + i]);
      A[size_t(i) * n + j] = A[size_t(j) * n + i] = avg;
    }
  if (syev(n, A, w)) return 1;
  // X = L^-T V: solve L^T X = V by back substitution (rows from the bottom)
  for (int i = n - 1; i >= 0; --i) {
    double *xi = &A[size_t(i) * n];
    for (int k = i + 1; k < n; ++k) {
      const double l = L[size_t(k) * n + i];
      const double *xk = &A[size_t(k) * n];
      for (int c = 0; c < n; ++c) xi[c] -= l * xk[c];
    }
    const double inv = 1.0 / L[size_t(i) * n + i];
    for (int c = 0; c < n; ++c) xi[c] *= inv;
  }
  return 0;
}

int chol_inv_t(int n, const double *S, double *X) {
  // S = U^T U with U = L^T upper (row-major), right-looking: row j of U, then the trailing update
  // A[i][i..] -= U[j][i] U[j][i..], all on contiguous row segments (vectorised axpys)
  std::vector<double> U(S, S + size_t(n) * n);
  for (int j = 0; j < n; ++j) {
    double *Uj = U.data() + size_t(j) * n;
    if (Uj[j] <= 0.0) return 1;
    const double d = std::sqrt(Uj[j]), inv = 1.0 / d;
    Uj[j] = d;
    for (int k = j + 1; k < n; ++k) Uj[k] *= inv;
    for (int i = j + 1; i < n; ++i) {
      double *__restrict__ Ai = U.data() + size_t(i) * n;
      const double *__restrict__ uj = Uj;
      const double f = uj[i];
      for (int k = i; k < n; ++k) Ai[k] -= f * uj[k];
    }
  }
  // X = L^-T = U^-1 (upper), by rows from the bottom: X[i][i..] = (e_i - sum_{k>i} U[i][k] X[k][i..]) / U[i][i]
  for (int i = n - 1; i >= 0; --i) {
    double *__restrict__ Xi = X + size_t(i) * n;
    const double *Ui = U.data() + size_t(i) * n;
    for (int c = 0; c < n; ++c) Xi[c] = c == i ? 1.0 : 0.0;
    for (int k = i + 1; k < n; ++k) {
      const double f = Ui[k];
      const double *__restrict__ Xk = X + size_t(k) * n;
      for (int c = k; c < n; ++c) Xi[c] -= f * Xk[c];
    }
    const double inv = 1.0 / Ui[i];
    for (int c = i; c < n; ++c) Xi[c] *= inv;
  }
  return 0;
}

int getrf(int n, double *A, int *piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(A[size_t(k) * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[size_t(i) * n + k]) > best) best = std::fabs(A[size_t(i) * n + k]), p = i;
    piv[k] = p;
    if (best == 0.0) return 1;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(A[size_t(k) * n + j], A[size_t(p) * n + j]);
    for (int i = k + 1; i < n; ++i) {
      const double f = A[size_t(i) * n + k] / A[size_t(k) * n + k];
      A[size_t(i) * n + k] = f;
      for (int j = k + 1; j < n; ++j) A[size_t(i) * n + j] -= f * A[size_t(k) * n + j];
    }
  }
  return 0;
}

void getrs(int n, const double *LU, const int *piv, double *b) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) std::swap(b[k], b[piv[k]]);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) b[i] -= LU[size_t(i) * n + j] * b[j];
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) b[i] -= LU[size_t(i) * n + j] * b[j];
    b[i] /= LU[size_t(i) * n + i];
  }
}

}  // namespace zq
