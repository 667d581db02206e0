/*
 * oracle.c -- CPU FP64 restatement of the reference's J/K, SCF, AO->MO transform and MP/CI consumers.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the zinq_amd HIP path: it may be
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing else.
 * The product (zinq_amd) never links, loads or calls it.
 *
 * Every function follows one reference function; loop orders and the order of floating-point
 * operations are kept (the reference computes with separate multiply and add through Value(T),
 * src/value.zig:37-47,127-137), and this file is compiled with -ffp-contract=off so no FMA is formed.
 * Library calls of the reference (OpenBLAS dgemm/ddot/dgetrf/dgetrs, LAPACK dsygvd/dsyevd) are replaced
 * by textbook loops: a partial-pivoting LU, a Cholesky reduction and a cyclic Jacobi eigensolver.
 *
 * Layouts: row-major, Matrix.at(i,j) = data[i*ncol+j] (src/tensor.zig:40);
 * Tensor(T,4).at({a,b,c,d}) = data[((a*n+b)*n+c)*n+d] (src/tensor.zig:274-288).
 * Dense ERIs g[p,q,r,s] = <pq|rs> = (pr|qs) as written by src/libint.cpp:143-150.
 */
#include <math.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <string.h>

#define AT2(A, n, i, j) ((A)[(size_t)(i) * (n) + (j)])
#define AT4(A, n, a, b, c, d) ((A)[(((size_t)(a) * (n) + (b)) * (n) + (c)) * (n) + (d)])

/* ------------------------------------------------------------------ linear algebra helpers */

/* LU with partial pivoting (what dgetrf computes); returns 0 on success, 1 on an exactly zero pivot
 * (dgetrf info > 0), as used through src/linear_algebra.zig:203-231. */
static int lu_factor(int n, double *A, int *piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(AT2(A, n, k, k));
    for (int i = k + 1; i < n; ++i)
      if (fabs(AT2(A, n, i, k)) > best) best = fabs(AT2(A, n, i, k)), p = i;
    piv[k] = p;
    if (best == 0.0) return 1;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double t = AT2(A, n, k, j);
        AT2(A, n, k, j) = AT2(A, n, p, j);
        AT2(A, n, p, j) = t;
      }
    for (int i = k + 1; i < n; ++i) {
      double f = AT2(A, n, i, k) / AT2(A, n, k, k);
      AT2(A, n, i, k) = f;
      for (int j = k + 1; j < n; ++j) AT2(A, n, i, j) -= f * AT2(A, n, k, j);
    }
  }
  return 0;
}

static void lu_solve(int n, const double *LU, const int *piv, double *b) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) {
      double t = b[k];
      b[k] = b[piv[k]];
      b[piv[k]] = t;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) b[i] -= AT2(LU, n, i, j) * b[j];
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) b[i] -= AT2(LU, n, i, j) * b[j];
    b[i] /= AT2(LU, n, i, i);
  }
}

/* Cyclic Jacobi eigensolver for a symmetric matrix A (destroyed).  Eigenvalues ascending in e,
 * eigenvectors in the columns of V (V[row*n + col]), matching dsyevd's output convention. */
int orc_eigh(int n, double *A, double *e, double *V) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) AT2(V, n, i, j) = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        tot += AT2(A, n, i, j) * AT2(A, n, i, j);
        if (i != j) off += AT2(A, n, i, j) * AT2(A, n, i, j);
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = AT2(A, n, p, q);
        if (apq == 0.0) continue;
        double app = AT2(A, n, p, p), aqq = AT2(A, n, q, q);
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = AT2(A, n, k, p), akq = AT2(A, n, k, q);
          AT2(A, n, k, p) = c * akp - s * akq;
          AT2(A, n, k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = AT2(A, n, p, k), aqk = AT2(A, n, q, k);
          AT2(A, n, p, k) = c * apk - s * aqk;
          AT2(A, n, q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = AT2(V, n, k, p), vkq = AT2(V, n, k, q);
          AT2(V, n, k, p) = c * vkp - s * vkq;
          AT2(V, n, k, q) = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) e[i] = AT2(A, n, i, i);
  /* selection sort ascending, permuting columns of V */
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j)
      if (e[j] < e[m]) m = j;
    if (m != i) {
      double t = e[i];
      e[i] = e[m], e[m] = t;
      for (int k = 0; k < n; ++k) {
        double v = AT2(V, n, k, i);
        AT2(V, n, k, i) = AT2(V, n, k, m);
        AT2(V, n, k, m) = v;
      }
    }
  }
  return 0;
}

/* Generalized symmetric-definite problem A C = B C e (dsygvd itype 1, src/linear_algebra.zig:184-200):
 * Cholesky B = L L^T, A' = L^-1 A L^-T, Jacobi on A', C = L^-T V.  Returns 1 if B is not positive
 * definite (LapackError). */
int orc_geigh(int n, const double *A, const double *B, double *e, double *C) {
  double *L = calloc((size_t)n * n, sizeof(double));
  double *W = malloc(sizeof(double) * n * n);
  double *V = malloc(sizeof(double) * n * n);
  int rc = 0;
  for (int j = 0; j < n && !rc; ++j) {
    double s = AT2(B, n, j, j);
    for (int k = 0; k < j; ++k) s -= AT2(L, n, j, k) * AT2(L, n, j, k);
    if (s <= 0.0) {
      rc = 1;
      break;
    }
    AT2(L, n, j, j) = sqrt(s);
    for (int i = j + 1; i < n; ++i) {
      double t = AT2(B, n, i, j);
      for (int k = 0; k < j; ++k) t -= AT2(L, n, i, k) * AT2(L, n, j, k);
      AT2(L, n, i, j) = t / AT2(L, n, j, j);
    }
  }
  if (!rc) {
    /* W = L^-1 A  (forward substitution per column), then A' = W L^-T = (L^-1 W^T)^T */
    for (int c = 0; c < n; ++c)
      for (int i = 0; i < n; ++i) {
        double t = AT2(A, n, i, c);
        for (int k = 0; k < i; ++k) t -= AT2(L, n, i, k) * AT2(W, n, k, c);
        AT2(W, n, i, c) = t / AT2(L, n, i, i);
      }
    double *Ap = malloc(sizeof(double) * n * n);
    for (int r = 0; r < n; ++r)
      for (int i = 0; i < n; ++i) {
        double t = AT2(W, n, r, i);
        for (int k = 0; k < i; ++k) t -= AT2(L, n, i, k) * AT2(Ap, n, r, k);
        AT2(Ap, n, r, i) = t / AT2(L, n, i, i);
      }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) {
        double avg = 0.5 * (AT2(Ap, n, i, j) + AT2(Ap, n, j, i));
        AT2(Ap, n, i, j) = AT2(Ap, n, j, i) = avg;
      }
    orc_eigh(n, Ap, e, V);
    /* C = L^-T V: back substitution L^T C = V */
    for (int c = 0; c < n; ++c)
      for (int i = n - 1; i >= 0; --i) {
        double t = AT2(V, n, i, c);
        for (int k = i + 1; k < n; ++k) t -= AT2(L, n, k, i) * AT2(C, n, k, c);
        AT2(C, n, i, c) = t / AT2(L, n, i, i);
      }
    free(Ap);
  }
  free(L), free(W), free(V);
  return rc;
}

/* ------------------------------------------------------------------ SCF pieces (src/hartree_fock.zig) */

/* getDensity, src/hartree_fock.zig:621-640.  Returns rms(P_new - P_old). */
double orc_get_density(int n, const double *C, int nocc, int generalized, double *P) {
  const double factor = generalized ? 1.0 : 2.0;
  double sum_sq_dp = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double sum = 0;
      for (int k = 0; k < nocc; ++k) sum += factor * AT2(C, n, i, k) * AT2(C, n, j, k);
      sum_sq_dp += (sum - AT2(P, n, i, j)) * (sum - AT2(P, n, i, j));
      AT2(P, n, i, j) = sum;
    }
  return sqrt(sum_sq_dp / (double)((size_t)n * n));
}

/* getEnergy (non-DFT branch), src/hartree_fock.zig:643-655. */
double orc_get_energy(int n, const double *H, const double *F, const double *P) {
  double energy = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) energy += 0.5 * AT2(P, n, i, j) * (AT2(H, n, i, j) + AT2(F, n, i, j));
  return energy;
}

/* getError, src/hartree_fock.zig:658-683: err = FPS - (FPS)^T (dgemm replaced by loops). */
void orc_get_error(int n, const double *F, const double *P, const double *S, double *err) {
  double *FP = calloc((size_t)n * n, sizeof(double)), *FPS = calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int k = 0; k < n; ++k) s += AT2(F, n, i, k) * AT2(P, n, k, j);
      AT2(FP, n, i, j) = s;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int k = 0; k < n; ++k) s += AT2(FP, n, i, k) * AT2(S, n, k, j);
      AT2(FPS, n, i, j) = s;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) AT2(err, n, i, j) = AT2(FPS, n, i, j) - AT2(FPS, n, j, i);
  free(FP), free(FPS);
}

/* getFock (non-DFT branch), src/hartree_fock.zig:686-735:
 * F = H; F[k,l] += P[i,j] (g[i,k,j,l] - x g[i,j,k,l]); F <- (F + F^T)/2, x = 1/2 (RHF) or 1 (GHF). */
void orc_get_fock(int n, const double *g, const double *H, const double *P, int generalized, double *F) {
  const double exch_factor = generalized ? 1.0 : 0.5;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) AT2(F, n, i, j) = AT2(H, n, i, j);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k)
        for (int l = 0; l < n; ++l)
          AT2(F, n, k, l) += AT2(P, n, i, j) * (AT4(g, n, i, k, j, l) - exch_factor * AT4(g, n, i, j, k, l));
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double avg = (AT2(F, n, i, j) + AT2(F, n, j, i)) / 2;
      AT2(F, n, i, j) = avg;
      AT2(F, n, j, i) = avg;
    }
}

/* getFock, DFT branch (src/hartree_fock.zig:701-719 and the symmetrisation :729-734): F = H; the Coulomb
 * loop F[l,m] += P[j,k] g[j,l,k,m]; if exx > 0 the exchange loop F[k,l] -= x exx P[i,j] g[i,j,k,l]
 * (x = 1/2 RHF, 1 GHF); then F += Vxc (the exchange-correlation potential the caller evaluated, the
 * pot.Vxc of :713-717; NULL = 0). */
void orc_get_fock_dft(int n, const double *g, const double *H, const double *P, int generalized, double exx,
                      const double *Vxc, double *F) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) AT2(F, n, i, j) = AT2(H, n, i, j);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k)
      for (int l = 0; l < n; ++l)
        for (int m = 0; m < n; ++m) AT2(F, n, l, m) += AT2(P, n, j, k) * AT4(g, n, j, l, k, m);
  if (exx > 0) {
    const double factor = generalized ? 1.0 : 0.5;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k)
          for (int l = 0; l < n; ++l) AT2(F, n, k, l) -= factor * exx * AT2(P, n, i, j) * AT4(g, n, i, j, k, l);
  }
  if (Vxc)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) AT2(F, n, j, k) += AT2(Vxc, n, j, k);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double avg = (AT2(F, n, i, j) + AT2(F, n, j, i)) / 2;
      AT2(F, n, i, j) = avg;
      AT2(F, n, j, i) = avg;
    }
}

/* The two-electron part of the HF nuclear gradient, src/hartree_fock.zig:263-273: for each coordinate
 * x < ncoord of the derivative tensor dg (ncoord dense n^4 blocks, the layout of ints.dg),
 * out[x] = sum_pqrs 0.5 P[p,q] P[r,s] (dg[x,p,r,q,s] - exx x_f dg[x,p,q,r,s]), x_f = 1/2 RHF, 1 GHF. */
void orc_grad_eri(int n, int ncoord, const double *dg, const double *P, int generalized, double exx, double *out) {
  const double exch_factor = generalized ? 1.0 : 0.5;
  const size_t n4 = (size_t)n * n * n * n;
  for (int x = 0; x < ncoord; ++x) {
    const double *d = dg + (size_t)x * n4;
    double g_G = 0;
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q)
        for (int r = 0; r < n; ++r)
          for (int s = 0; s < n; ++s) {
            const double dg1 = AT4(d, n, p, r, q, s), dg2 = AT4(d, n, p, q, r, s);
            g_G += 0.5 * AT2(P, n, p, q) * AT2(P, n, r, s) * (dg1 - exx * exch_factor * dg2);
          }
    out[x] = g_G;
  }
}

/* diis, src/hartree_fock.zig:176-227.  fh/eh: m matrices of n*n.  Returns 1 if the LU solve fails. */
int orc_diis(int m, int n, const double *fh, const double *eh, double *F) {
  if (m < 2) return 0;
  const int d = m + 1;
  double *B = calloc((size_t)d * d, sizeof(double)), *c = calloc(d, sizeof(double));
  int *piv = malloc(sizeof(int) * d);
  const size_t nn = (size_t)n * n;
  for (int i = 0; i < m; ++i)
    for (int j = i; j < m; ++j) {
      double sum = 0;
      for (size_t k = 0; k < nn; ++k) sum += eh[i * nn + k] * eh[j * nn + k];
      AT2(B, d, i, j) = sum;
      AT2(B, d, j, i) = sum;
    }
  for (int i = 0; i < m; ++i) AT2(B, d, i, m) = -1.0, AT2(B, d, m, i) = -1.0;
  c[m] = -1.0;
  int rc = lu_factor(d, B, piv);
  if (!rc) {
    lu_solve(d, B, piv, c);
    memset(F, 0, sizeof(double) * nn);
    for (int i = 0; i < m; ++i)
      for (size_t k = 0; k < nn; ++k) F[k] += c[i] * fh[i * nn + k];
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        double avg = (AT2(F, n, i, j) + AT2(F, n, j, i)) / 2;
        AT2(F, n, i, j) = avg;
        AT2(F, n, j, i) = avg;
      }
  }
  free(B), free(c), free(piv);
  return rc;
}

/* scf, src/hartree_fock.zig:738-875, preceded by the core-Hamiltonian guess of :425-431.
 * Inputs are the (spin-blocked for GHF) H, S, g of dimension n.  diis <= 0 disables DIIS (diis = null).
 * Outputs: C, e, P, F (the last diagonalised, DIIS-extrapolated F); returns 0, or 1 for
 * ScfDidNotConverge, 2 for LapackError.  *iters_out receives the number of iterations run. */
void orc_get_fock_packed(int n, const double *pk, const double *H, const double *P, double x, double *F);
void orc_get_fock_par(int n, const double *g, const double *H, const double *P, int generalized, double *F);

static int scf_impl(int n, int nocc, int generalized, const double *H, const double *S, const double *g, int packed,
                    double VN, int diis, int iterations, double threshold, double *E_out, double *C, double *e,
                    double *P, double *F, int *iters_out) {
  const size_t nn = (size_t)n * n;
  memset(P, 0, sizeof(double) * nn);
  if (orc_geigh(n, H, S, e, C)) return 2;
  orc_get_density(n, C, nocc, generalized, P);
  const int maxh = diis > 0 ? diis : 1;
  double *fh = malloc(sizeof(double) * nn * maxh), *eh = malloc(sizeof(double) * nn * maxh);
  int m = 0;
  double e_old = VN, e_new = VN;
  int converged = 0, it;
  for (it = 0; it < iterations; ++it) {
    if (packed == 1)
      orc_get_fock_packed(n, g, H, P, generalized ? 1.0 : 0.5, F);
    else if (packed == 2)
      orc_get_fock_par(n, g, H, P, generalized, F);
    else
      orc_get_fock(n, g, H, P, generalized, F);
    e_new = orc_get_energy(n, H, F, P) + VN;
    if (diis > 0) {
      if (m >= diis) {
        memmove(fh, fh + nn, sizeof(double) * nn * (m - 1));
        memmove(eh, eh + nn, sizeof(double) * nn * (m - 1));
        --m;
      }
      memcpy(fh + nn * m, F, sizeof(double) * nn);
      orc_get_error(n, F, P, S, eh + nn * m);
      ++m;
      if (orc_diis(m, n, fh, eh, F)) {
        /* reset the history to the current (non-extrapolated) entry, :821-845 */
        memcpy(fh, F, sizeof(double) * nn);
        orc_get_error(n, F, P, S, eh);
        m = 1;
      }
    }
    if (orc_geigh(n, F, S, e, C)) {
      free(fh), free(eh);
      return 2;
    }
    const double delta_energy = fabs(e_new - e_old);
    const double p_rm = orc_get_density(n, C, nocc, generalized, P);
    e_old = e_new;
    if (it > 0 && delta_energy < threshold && p_rm < threshold) {
      converged = 1;
      ++it;
      break;
    }
  }
  free(fh), free(eh);
  *E_out = e_new;
  *iters_out = it;
  return converged ? 0 : 1;
}

int orc_scf(int n, int nocc, int generalized, const double *H, const double *S, const double *g, double VN, int diis,
            int iterations, double threshold, double *E_out, double *C, double *e, double *P, double *F,
            int *iters_out) {
  return scf_impl(n, nocc, generalized, H, S, g, 0, VN, diis, iterations, threshold, E_out, C, e, P, F, iters_out);
}

/* orc_scf with the Fock builds on all cores (orc_get_fock_par, bitwise orc_get_fock): bitwise orc_scf. */
int orc_scf_par(int n, int nocc, int generalized, const double *H, const double *S, const double *g, double VN, int diis,
                int iterations, double threshold, double *E_out, double *C, double *e, double *P, double *F,
                int *iters_out) {
  return scf_impl(n, nocc, generalized, H, S, g, 2, VN, diis, iterations, threshold, E_out, C, e, P, F, iters_out);
}

/* orc_scf with the Fock build read from canonical packed ERIs (orc_get_fock_packed): bitwise orc_scf, for
 * sizes whose dense tensor does not fit. */
int orc_scf_packed(int n, int nocc, int generalized, const double *H, const double *S, const double *pk, double VN,
                   int diis, int iterations, double threshold, double *E_out, double *C, double *e, double *P,
                   double *F, int *iters_out) {
  return scf_impl(n, nocc, generalized, H, S, pk, 1, VN, diis, iterations, threshold, E_out, C, e, P, F, iters_out);
}

/* ------------------------------------------------------------------ transforms (src/integral_transform.zig) */

/* ao2mo_pp, src/integral_transform.zig:12-34. */
void orc_ao2mo_pp(int N, const double *A, const double *C, double *out) {
  for (int p = 0; p < N; ++p)
    for (int q = 0; q < N; ++q) {
      double sum = 0;
      for (int mu = 0; mu < N; ++mu)
        for (int nu = 0; nu < N; ++nu) sum = sum + AT2(C, N, mu, p) * AT2(C, N, nu, q) * AT2(A, N, mu, nu);
      AT2(out, N, p, q) = sum;
    }
}

/* ao2mo_pppp, src/integral_transform.zig:37-79: four quarter loops, physicist order kept. */
void orc_ao2mo_pppp(int N, const double *g, const double *C, double *out) {
  const size_t n4 = (size_t)N * N * N * N;
  double *t1 = calloc(n4, sizeof(double)), *t2 = calloc(n4, sizeof(double)), *t3 = calloc(n4, sizeof(double));
  memset(out, 0, sizeof(double) * n4);
  for (int i = 0; i < N; ++i)
    for (int lam = 0; lam < N; ++lam)
      for (int nu = 0; nu < N; ++nu)
        for (int mu = 0; mu < N; ++mu)
          for (int sg = 0; sg < N; ++sg) AT4(t1, N, i, lam, nu, sg) += AT2(C, N, mu, i) * AT4(g, N, mu, lam, nu, sg);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int nu = 0; nu < N; ++nu)
        for (int lam = 0; lam < N; ++lam)
          for (int sg = 0; sg < N; ++sg) AT4(t2, N, i, j, nu, sg) += AT2(C, N, lam, j) * AT4(t1, N, i, lam, nu, sg);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int a = 0; a < N; ++a)
        for (int nu = 0; nu < N; ++nu)
          for (int sg = 0; sg < N; ++sg) AT4(t3, N, i, j, a, sg) += AT2(C, N, nu, a) * AT4(t2, N, i, j, nu, sg);
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int a = 0; a < N; ++a)
        for (int sg = 0; sg < N; ++sg)
          for (int b = 0; b < N; ++b) AT4(out, N, i, j, a, b) += AT2(C, N, sg, b) * AT4(t3, N, i, j, a, sg);
  free(t1), free(t2), free(t3);
}

/* ao2mo_pppp over ScalarDual values (src/integral_transform.zig:37-79 with T = ScalarDual, as
 * src/moller_plesset.zig:304-333 runs it): (g, dg) and (C, dC) are the value and derivative parts;
 * every += a * b is the dual product (val a.val b.val, der a.val b.der + b.val a.der, src/dual.zig:51-53)
 * and the dual sum. */
void orc_ao2mo_pppp_dual(int N, const double *g, const double *dg, const double *C, const double *dC, double *out,
                         double *dout) {
  const size_t n4 = (size_t)N * N * N * N;
  double *t1 = calloc(n4, sizeof(double)), *t2 = calloc(n4, sizeof(double)), *t3 = calloc(n4, sizeof(double));
  double *d1 = calloc(n4, sizeof(double)), *d2 = calloc(n4, sizeof(double)), *d3 = calloc(n4, sizeof(double));
  memset(out, 0, sizeof(double) * n4);
  memset(dout, 0, sizeof(double) * n4);
  for (int i = 0; i < N; ++i)
    for (int lam = 0; lam < N; ++lam)
      for (int nu = 0; nu < N; ++nu)
        for (int mu = 0; mu < N; ++mu)
          for (int sg = 0; sg < N; ++sg) {
            const double a = AT2(C, N, mu, i), da = AT2(dC, N, mu, i), b = AT4(g, N, mu, lam, nu, sg),
                         db = AT4(dg, N, mu, lam, nu, sg);
            AT4(t1, N, i, lam, nu, sg) += a * b;
            AT4(d1, N, i, lam, nu, sg) += a * db + b * da;
          }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int nu = 0; nu < N; ++nu)
        for (int lam = 0; lam < N; ++lam)
          for (int sg = 0; sg < N; ++sg) {
            const double a = AT2(C, N, lam, j), da = AT2(dC, N, lam, j), b = AT4(t1, N, i, lam, nu, sg),
                         db = AT4(d1, N, i, lam, nu, sg);
            AT4(t2, N, i, j, nu, sg) += a * b;
            AT4(d2, N, i, j, nu, sg) += a * db + b * da;
          }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int a_ = 0; a_ < N; ++a_)
        for (int nu = 0; nu < N; ++nu)
          for (int sg = 0; sg < N; ++sg) {
            const double a = AT2(C, N, nu, a_), da = AT2(dC, N, nu, a_), b = AT4(t2, N, i, j, nu, sg),
                         db = AT4(d2, N, i, j, nu, sg);
            AT4(t3, N, i, j, a_, sg) += a * b;
            AT4(d3, N, i, j, a_, sg) += a * db + b * da;
          }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j)
      for (int a_ = 0; a_ < N; ++a_)
        for (int sg = 0; sg < N; ++sg)
          for (int b_ = 0; b_ < N; ++b_) {
            const double a = AT2(C, N, sg, b_), da = AT2(dC, N, sg, b_), b = AT4(t3, N, i, j, a_, sg),
                         db = AT4(d3, N, i, j, a_, sg);
            AT4(out, N, i, j, a_, b_) += a * b;
            AT4(dout, N, i, j, a_, b_) += a * db + b * da;
          }
  free(t1), free(t2), free(t3), free(d1), free(d2), free(d3);
}

/* ao2mo_oovv, src/integral_transform.zig:165-219: out[i,j,a,b] = <ij|ab> = (ia|jb), i,j occupied,
 * a,b virtual, shape (nocc, nocc, nvir, nvir). */
void orc_ao2mo_oovv(int N, int nocc, const double *g, const double *C, double *out) {
  const int nvir = N - nocc;
  const size_t o = nocc, v = nvir, n = N;
  double *t1 = calloc(o * n * n * n, sizeof(double)), *t2 = calloc(o * o * n * n, sizeof(double));
  double *t3 = calloc(o * o * v * n, sizeof(double));
  memset(out, 0, sizeof(double) * o * o * v * v);
#define T1(i, l, nu, s) t1[(((i)*n + (l)) * n + (nu)) * n + (s)]
#define T2(i, j, nu, s) t2[(((i)*o + (j)) * n + (nu)) * n + (s)]
#define T3(i, j, a, s) t3[(((i)*o + (j)) * v + (a)) * n + (s)]
#define OUT(i, j, a, b) out[(((i)*o + (j)) * v + (a)) * v + (b)]
  for (size_t i = 0; i < o; ++i)
    for (size_t lam = 0; lam < n; ++lam)
      for (size_t nu = 0; nu < n; ++nu)
        for (size_t mu = 0; mu < n; ++mu)
          for (size_t sg = 0; sg < n; ++sg) T1(i, lam, nu, sg) += AT2(C, N, mu, i) * AT4(g, N, mu, lam, nu, sg);
  for (size_t i = 0; i < o; ++i)
    for (size_t j = 0; j < o; ++j)
      for (size_t nu = 0; nu < n; ++nu)
        for (size_t lam = 0; lam < n; ++lam)
          for (size_t sg = 0; sg < n; ++sg) T2(i, j, nu, sg) += AT2(C, N, lam, j) * T1(i, lam, nu, sg);
  for (size_t i = 0; i < o; ++i)
    for (size_t j = 0; j < o; ++j)
      for (size_t a = o; a < n; ++a)
        for (size_t nu = 0; nu < n; ++nu)
          for (size_t sg = 0; sg < n; ++sg) T3(i, j, a - o, sg) += AT2(C, N, nu, a) * T2(i, j, nu, sg);
  for (size_t i = 0; i < o; ++i)
    for (size_t j = 0; j < o; ++j)
      for (size_t a = 0; a < v; ++a)
        for (size_t sg = 0; sg < n; ++sg)
          for (size_t b = o; b < n; ++b) OUT(i, j, a, b - o) += AT2(C, N, sg, b) * T3(i, j, a, sg);
#undef T1
#undef T2
#undef T3
#undef OUT
  free(t1), free(t2), free(t3);
}

/* The same arithmetic on all cores (OpenMP), for the CPU baseline's many-core mode: every output element
 * is accumulated in the faithful functions' order, so the results are bitwise those of orc_get_fock and
 * orc_ao2mo_oovv; only independent outputs are split across threads. */
void orc_get_fock_par(int n, const double *g, const double *H, const double *P, int generalized, double *F) {
  const double exch_factor = generalized ? 1.0 : 0.5;
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < n; ++k)
    for (int l = 0; l < n; ++l) {
      double f = AT2(H, n, k, l);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) f += AT2(P, n, i, j) * (AT4(g, n, i, k, j, l) - exch_factor * AT4(g, n, i, j, k, l));
      AT2(F, n, k, l) = f;
    }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double avg = (AT2(F, n, i, j) + AT2(F, n, j, i)) / 2;
      AT2(F, n, i, j) = avg;
      AT2(F, n, j, i) = avg;
    }
}

void orc_ao2mo_oovv_par(int N, int nocc, const double *g, const double *C, double *out) {
  const int nvir = N - nocc;
  const size_t o = nocc, v = nvir, n = N;
  double *t1 = calloc(o * n * n * n, sizeof(double)), *t2 = calloc(o * o * n * n, sizeof(double));
  double *t3 = calloc(o * o * v * n, sizeof(double));
  memset(out, 0, sizeof(double) * o * o * v * v);
#define T1(i, l, nu, s) t1[(((i)*n + (l)) * n + (nu)) * n + (s)]
#define T2(i, j, nu, s) t2[(((i)*o + (j)) * n + (nu)) * n + (s)]
#define T3(i, j, a, s) t3[(((i)*o + (j)) * v + (a)) * n + (s)]
#define OUT(i, j, a, b) out[(((i)*o + (j)) * v + (a)) * v + (b)]
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (size_t i = 0; i < o; ++i)
    for (size_t lam = 0; lam < n; ++lam)
      for (size_t nu = 0; nu < n; ++nu)
        for (size_t mu = 0; mu < n; ++mu)
          for (size_t sg = 0; sg < n; ++sg) T1(i, lam, nu, sg) += AT2(C, N, mu, i) * AT4(g, N, mu, lam, nu, sg);
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (size_t i = 0; i < o; ++i)
    for (size_t j = 0; j < o; ++j)
      for (size_t nu = 0; nu < n; ++nu)
        for (size_t lam = 0; lam < n; ++lam)
          for (size_t sg = 0; sg < n; ++sg) T2(i, j, nu, sg) += AT2(C, N, lam, j) * T1(i, lam, nu, sg);
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (size_t i = 0; i < o; ++i)
    for (size_t j = 0; j < o; ++j)
      for (size_t a = o; a < n; ++a)
        for (size_t nu = 0; nu < n; ++nu)
          for (size_t sg = 0; sg < n; ++sg) T3(i, j, a - o, sg) += AT2(C, N, nu, a) * T2(i, j, nu, sg);
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (size_t i = 0; i < o; ++i)
    for (size_t j = 0; j < o; ++j)
      for (size_t a = 0; a < v; ++a)
        for (size_t sg = 0; sg < n; ++sg)
          for (size_t b = o; b < n; ++b) OUT(i, j, a, b - o) += AT2(C, N, sg, b) * T3(i, j, a, sg);
#undef T1
#undef T2
#undef T3
#undef OUT
  free(t1), free(t2), free(t3);
}

/* ------------------------------------------------------------------ the same loops over packed ERIs
 * For sizes whose dense n^4 tensor does not fit (n = 192: 10.9 GB, n = 240: 26.5 GB) the loops above are
 * restated reading the canonical 8-fold packed array the integral engine writes (zinq_ints.h:
 * out[I(I+1)/2+J] = (pq|rs), I = p(p+1)/2+q >= J = r(r+1)/2+s).  Each value fetched is the one the dense
 * tensor holds at that position (the dense tensor is scattered from the same unique integrals,
 * src/libint.cpp:143-150) and the sums run in the same order, so the results are bitwise those of
 * orc_get_fock / orc_ao2mo_oovv on the rows or pairs computed. */
static inline size_t pk_pair(size_t p, size_t q) { return p >= q ? p * (p + 1) / 2 + q : q * (q + 1) / 2 + p; }

/* chemist (pq|rs) = dense g[p,r,q,s] */
static inline double pk_chem(const double *pk, size_t p, size_t q, size_t r, size_t s) {
  const size_t I = pk_pair(p, q), J = pk_pair(r, s);
  return I >= J ? pk[I * (I + 1) / 2 + J] : pk[J * (J + 1) / 2 + I];
}

/* Rows `rows[0..nrows)` of getFock's symmetrised F (src/hartree_fock.zig:686-735, non-DFT branch) with an
 * explicit exchange factor x: F[k,l] = H[k,l] + sum_ij P[i,j] (g[i,k,j,l] - x g[i,j,k,l]), row k of
 * (F + F^T)/2.  Frows is nrows x n. */
void orc_fock_rows_packed(int n, const double *pk, const double *H, const double *P, double x, int nrows,
                          const int *rows, double *Frows) {
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int r = 0; r < nrows; ++r)
    for (int l = 0; l < n; ++l) {
      const int k = rows[r];
      double fkl = AT2(H, n, k, l), flk = AT2(H, n, l, k);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          /* g[i,k,j,l] = (ij|kl), g[i,j,k,l] = (ik|jl) */
          fkl += AT2(P, n, i, j) * (pk_chem(pk, i, j, k, l) - x * pk_chem(pk, i, k, j, l));
          flk += AT2(P, n, i, j) * (pk_chem(pk, i, j, l, k) - x * pk_chem(pk, i, l, j, k));
        }
      AT2(Frows, n, r, l) = k == l ? fkl : (fkl + flk) / 2;
    }
}

/* The whole of getFock (src/hartree_fock.zig:686-735, non-DFT branch) from packed ERIs, exchange factor x:
 * each F[k,l] accumulated over i, j ascending from H[k,l], then (F + F^T)/2 as orc_get_fock does. */
void orc_get_fock_packed(int n, const double *pk, const double *H, const double *P, double x, double *F) {
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < n; ++k)
    for (int l = 0; l < n; ++l) {
      double f = AT2(H, n, k, l);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) f += AT2(P, n, i, j) * (pk_chem(pk, i, j, k, l) - x * pk_chem(pk, i, k, j, l));
      AT2(F, n, k, l) = f;
    }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double avg = (AT2(F, n, i, j) + AT2(F, n, j, i)) / 2;
      AT2(F, n, i, j) = avg;
      AT2(F, n, j, i) = avg;
    }
}

/* (ia|jb) = <ij|ab> of ao2mo_oovv (src/integral_transform.zig:165-219) for the occupied pairs
 * (pi[t], pj[t]), t < npairs: out[t] is the v x v block [a][b].  The four quarter sums run over mu, lam,
 * nu, sg in ascending order as in orc_ao2mo_oovv; the first quarter is kept for one i at a time, indexed
 * by the canonical ket pair (lam sg) since g[mu,lam,nu,sg] = (mu nu|lam sg) is symmetric in it. */
void orc_ovov_pairs_packed(int N, int nocc, const double *pk, const double *C, int npairs, const int *pi,
                           const int *pj, double *out) {
  const size_t n = N, o = nocc, v = N - nocc, P = n * (n + 1) / 2;
  double *t1 = malloc(sizeof(double) * n * P); /* t1[nu][pair(lam,sg)] for the current i */
  double *t2 = malloc(sizeof(double) * n * n), *t3 = malloc(sizeof(double) * v * n);
  int cur = -1;
  for (int t = 0; t < npairs; ++t) {
    const size_t i = pi[t], j = pj[t];
    if (i >= o || j >= o) continue;
    if ((int)i != cur) {
      cur = (int)i;
      /* ket pair outermost: for I = (mu nu) below J the reads run along the packed row of J */
#pragma omp parallel for schedule(dynamic, 64)
      for (size_t J = 0; J < P; ++J) {
        size_t lam = 0;
        while ((lam + 1) * (lam + 2) / 2 <= J) ++lam;
        const size_t sg = J - lam * (lam + 1) / 2;
        for (size_t nu = 0; nu < n; ++nu) {
          double s = 0;
          for (size_t mu = 0; mu < n; ++mu) s += AT2(C, N, mu, i) * pk_chem(pk, mu, nu, lam, sg);
          t1[nu * P + J] = s;
        }
      }
    }
#pragma omp parallel for
    for (size_t nu = 0; nu < n; ++nu)
      for (size_t sg = 0; sg < n; ++sg) {
        double s = 0;
        for (size_t lam = 0; lam < n; ++lam) s += AT2(C, N, lam, j) * t1[nu * P + pk_pair(lam, sg)];
        t2[nu * n + sg] = s;
      }
#pragma omp parallel for
    for (size_t a = 0; a < v; ++a)
      for (size_t sg = 0; sg < n; ++sg) {
        double s = 0;
        for (size_t nu = 0; nu < n; ++nu) s += AT2(C, N, nu, o + a) * t2[nu * n + sg];
        t3[a * n + sg] = s;
      }
#pragma omp parallel for
    for (size_t a = 0; a < v; ++a)
      for (size_t b = 0; b < v; ++b) {
        double s = 0;
        for (size_t sg = 0; sg < n; ++sg) s += AT2(C, N, sg, o + b) * t3[a * n + sg];
        out[((size_t)t * v + a) * v + b] = s;
      }
  }
  free(t1), free(t2), free(t3);
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ao2so_coef, src/integral_transform.zig:82-101. */
void orc_ao2so_coef(int n, const double *C, double *Cso) {
  const int N = 2 * n;
  memset(Cso, 0, sizeof(double) * N * N);
  for (int i = 0; i < n; ++i)
    for (int mu = 0; mu < n; ++mu) {
      AT2(Cso, N, mu, 2 * i + 0) = AT2(C, n, mu, i);
      AT2(Cso, N, mu + n, 2 * i + 1) = AT2(C, n, mu, i);
    }
}

/* ao2so_pp, src/integral_transform.zig:104-117 (also the spin-blocked one-electron matrices of
 * src/molecular_system.zig kineticSpin/nuclearSpin/overlapSpin). */
void orc_ao2so_pp(int n, const double *A, double *Aso) {
  const int N = 2 * n;
  memset(Aso, 0, sizeof(double) * N * N);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      AT2(Aso, N, i, j) = AT2(A, n, i, j);
      AT2(Aso, N, i + n, j + n) = AT2(A, n, i, j);
    }
}

/* ao2so_pppp, src/integral_transform.zig:120-142 (identical to coulombSpin,
 * src/molecular_system.zig:161-182). */
void orc_ao2so_pppp(int n, const double *g, double *gso) {
  const int N = 2 * n;
  memset(gso, 0, sizeof(double) * (size_t)N * N * N * N);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k)
        for (int l = 0; l < n; ++l) {
          double val = AT4(g, n, i, j, k, l);
          AT4(gso, N, i, j, k, l) = val;
          AT4(gso, N, i + n, j, k + n, l) = val;
          AT4(gso, N, i, j + n, k, l + n) = val;
          AT4(gso, N, i + n, j + n, k + n, l + n) = val;
        }
}

/* ------------------------------------------------------------------ determinants (configuration_interaction.zig) */

typedef struct {
  int ndet, nel;
  int *occ; /* ndet * nel, ascending */
} Dets;

/* generateCombinations, src/configuration_interaction.zig:462-510: lexicographic k-subsets of
 * {0..n-1} shifted by offset. */
static int gen_combs(int n, int k, int offset, int **out) {
  if (k > n) {
    *out = NULL;
    return 0;
  }
  int cnt = 1;
  for (int i = 0; i < k; ++i) cnt = cnt * (n - i) / (i + 1);
  int *r = malloc(sizeof(int) * (k > 0 ? k : 1) * cnt), cur[64];
  for (int i = 0; i < k; ++i) cur[i] = i;
  for (int c = 0; c < cnt; ++c) {
    for (int i = 0; i < k; ++i) r[c * k + i] = cur[i] + offset;
    int i = k - 1;
    while (i >= 0 && cur[i] == n - k + i) --i;
    if (i < 0) break;
    cur[i]++;
    for (int j = i + 1; j < k; ++j) cur[j] = cur[j - 1] + 1;
  }
  *out = r;
  return cnt;
}

static int cmp_int(const void *a, const void *b) { return *(const int *)a - *(const int *)b; }

/* generateDets, src/configuration_interaction.zig:126-183. */
static Dets gen_dets(int nel, int nsp, const int *exc, int nexc) {
  Dets d;
  d.nel = nel;
  size_t cap = 1;
  for (int x = 0; x < nexc; ++x) {
    int k = exc[x];
    if (k == 0 || k > nel || k > nsp - nel) continue;
    size_t a = 1, b = 1;
    for (int i = 0; i < k; ++i) a = a * (nel - i) / (i + 1);
    for (int i = 0; i < k; ++i) b = b * (nsp - nel - i) / (i + 1);
    cap += a * b;
  }
  d.occ = malloc(sizeof(int) * nel * cap);
  d.ndet = 0;
  for (int i = 0; i < nel; ++i) d.occ[i] = i;
  d.ndet = 1;
  for (int x = 0; x < nexc; ++x) {
    int k = exc[x];
    if (k == 0 || k > nel || k > nsp - nel) continue;
    int *oc, *vc;
    int no = gen_combs(nel, k, 0, &oc), nv = gen_combs(nsp - nel, k, nel, &vc);
    for (int j = 0; j < no; ++j)
      for (int l = 0; l < nv; ++l) {
        int *dst = d.occ + (size_t)d.ndet * nel, c = 0;
        for (int p = 0; p < nel; ++p) {
          int found = 0;
          for (int q = 0; q < k; ++q)
            if (oc[j * k + q] == p) found = 1;
          if (!found) dst[c++] = p;
        }
        for (int q = 0; q < k; ++q) dst[c++] = vc[l * k + q];
        qsort(dst, nel, sizeof(int), cmp_int);
        d.ndet++;
      }
    free(oc), free(vc);
  }
  return d;
}

/* signOfExcitations, src/configuration_interaction.zig:569-577. */
static int sign_exc(const int *S, int ns, const int *U, int nu) {
  size_t e = 0;
  for (int j = 0; j < nu; ++j) {
    int pos = 0;
    while (pos < ns && S[pos] != U[j]) ++pos;
    e += (size_t)(pos - j);
  }
  return (e % 2 == 0) ? 1 : -1;
}

/* slater, src/configuration_interaction.zig:323-432 (Slater-Condon rules). */
static double slater(const int *A, const int *B, int nel, int N, const double *H, const double *g) {
  int dA[2], dB[2], na = 0, nb = 0, ia = 0, ib = 0;
  while (ia < nel && ib < nel) {
    if (A[ia] < B[ib]) {
      if (na == 2) return 0.0;
      dA[na++] = A[ia++];
    } else if (A[ia] > B[ib]) {
      if (nb == 2) return 0.0;
      dB[nb++] = B[ib++];
    } else {
      ++ia, ++ib;
    }
  }
  for (; ia < nel; ++ia) {
    if (na == 2) return 0.0;
    dA[na++] = A[ia];
  }
  for (; ib < nel; ++ib) {
    if (nb == 2) return 0.0;
    dB[nb++] = B[ib];
  }
  const int sA = sign_exc(A, nel, dA, na), sB = sign_exc(B, nel, dB, nb);
  if (na == 0) {
    double sum = 0;
    for (int i = 0; i < nel; ++i) sum = sum + AT2(H, N, A[i], A[i]);
    for (int i = 0; i < nel; ++i)
      for (int j = i + 1; j < nel; ++j) sum = sum + (AT4(g, N, A[i], A[j], A[i], A[j]) - AT4(g, N, A[i], A[j], A[j], A[i]));
    return sum;
  }
  if (na == 1) {
    const int m = dA[0], p = dB[0];
    double term = AT2(H, N, m, p);
    for (int i = 0; i < nel; ++i)
      if (A[i] != m) term = term + (AT4(g, N, m, A[i], p, A[i]) - AT4(g, N, m, A[i], A[i], p));
    return term * (double)(sA * sB);
  }
  if (na == 2) {
    const int m = dA[0], nn = dA[1], p = dB[0], q = dB[1];
    return (AT4(g, N, m, nn, p, q) - AT4(g, N, m, nn, q, p)) * (double)(sA * sB);
  }
  return 0.0;
}

/* Spin-orbital MO integrals of the reference's consumers: RHF input is expanded with ao2so_coef /
 * ao2so_pppp first (generalized == 0); GHF input is already spin-blocked (generalized == 1). */
static void so_transform(int n, const double *g, const double *C, int generalized, double **g_MS_out, double **C_SO_out,
                         int *nsp_out) {
  const int nsp = generalized ? n : 2 * n;
  const size_t n4 = (size_t)nsp * nsp * nsp * nsp;
  double *C_SO = malloc(sizeof(double) * nsp * nsp);
  double *g_MS = malloc(sizeof(double) * n4);
  if (generalized) {
    memcpy(C_SO, C, sizeof(double) * nsp * nsp);
    orc_ao2mo_pppp(nsp, g, C_SO, g_MS);
  } else {
    double *g_SO = malloc(sizeof(double) * n4);
    orc_ao2so_coef(n, C, C_SO);
    orc_ao2so_pppp(n, g, g_SO);
    orc_ao2mo_pppp(nsp, g_SO, C_SO, g_MS);
    free(g_SO);
  }
  *g_MS_out = g_MS, *C_SO_out = C_SO, *nsp_out = nsp;
}

/* mp, src/moller_plesset.zig:340-481: Rayleigh-Schroedinger perturbation theory in the determinant
 * space of singles..order-fold excitations.  energies[0..order].  For order == 2 only row 0 of H_CI is
 * formed (the recursion reads nothing else at that order; the arithmetic is unchanged), which keeps
 * cc-pVDZ water (32016 determinants) in reach. */
int orc_mp(int order, int n, const double *g, const double *C, const double *e, int nocc, int generalized,
           double *energies) {
  int nsp;
  double *g_MS, *C_SO;
  so_transform(n, g, C, generalized, &g_MS, &C_SO, &nsp);
  const int nel = generalized ? nocc : 2 * nocc;
  double *e_SO = malloc(sizeof(double) * nsp);
  for (int p = 0; p < nsp; ++p) e_SO[p] = generalized ? e[p] : e[p / 2];
  double *H_MS = malloc(sizeof(double) * nsp * nsp);
  for (int p = 0; p < nsp; ++p)
    for (int q = 0; q < nsp; ++q) {
      double sum = 0;
      for (int i = 0; i < nel; ++i) sum = sum + (AT4(g_MS, nsp, p, i, q, i) - AT4(g_MS, nsp, p, i, i, q));
      AT2(H_MS, nsp, p, q) = ((p == q) ? e_SO[p] : 0.0) - sum;
    }
  int exc[16], nexc = order < nel ? order : nel;
  if (nsp - nel < nexc) nexc = nsp - nel;
  for (int i = 0; i < nexc; ++i) exc[i] = i + 1;
  Dets d = gen_dets(nel, nsp, exc, nexc);
  const size_t nd = d.ndet;
  double *E0 = malloc(sizeof(double) * nd);
  for (size_t i = 0; i < nd; ++i) {
    double s = 0;
    for (int k = 0; k < nel; ++k) s = s + e_SO[d.occ[i * nel + k]];
    E0[i] = s;
  }
  double *E = calloc(order + 1, sizeof(double));
  E[0] = E0[0];
  if (order == 2) {
    double *row = malloc(sizeof(double) * nd);
    for (size_t j = 0; j < nd; ++j) row[j] = slater(d.occ, d.occ + j * nel, nel, nsp, H_MS, g_MS);
    E[1] = row[0] - E0[0];
    double sum = 0;
    for (size_t j = 1; j < nd; ++j) sum = sum + row[j] * (row[j] / (E0[0] - E0[j]));
    E[2] = sum;
    free(row);
  } else {
    double *H = malloc(sizeof(double) * nd * nd);
    for (size_t i = 0; i < nd; ++i)
      for (size_t j = i; j < nd; ++j) {
        double v = slater(d.occ + i * nel, d.occ + j * nel, nel, nsp, H_MS, g_MS);
        H[i * nd + j] = v, H[j * nd + i] = v;
      }
    double *Cc = calloc((size_t)(order + 1) * nd, sizeof(double));
    Cc[0] = 1.0;
    for (int k = 1; k <= order; ++k) {
      if (k == 1) E[1] = H[0] - E0[0];
      if (k > 1) {
        double sum = 0;
        for (size_t j = 1; j < nd; ++j) sum = sum + H[j] * Cc[(size_t)(k - 1) * nd + j];
        E[k] = sum;
      }
      if (k == order) break;
      for (size_t i = 1; i < nd; ++i) {
        double term1 = 0, term3 = 0;
        for (size_t j = 0; j < nd; ++j) term1 = term1 + H[i * nd + j] * Cc[(size_t)(k - 1) * nd + j];
        double term2 = E0[i] * Cc[(size_t)(k - 1) * nd + i];
        for (int j = 1; j <= k; ++j) term3 = term3 + E[j] * Cc[(size_t)(k - j) * nd + i];
        Cc[(size_t)k * nd + i] = (term1 - term2 - term3) / (E0[0] - E0[i]);
      }
    }
    free(H), free(Cc);
  }
  for (int k = 0; k <= order; ++k) energies[k] = E[k];
  free(E), free(E0), free(d.occ), free(H_MS), free(e_SO), free(g_MS), free(C_SO);
  return 0;
}

/* Closed-shell MP2 correlation energy from (ia|jb) = <ij|ab> in ao2mo_oovv order; algebraically equal
 * to orc_mp's order-2 term for RHF references (the singles couple with exactly zero because H_MS is
 * built from the orbital energies, src/moller_plesset.zig:377-388). */
double orc_mp2_closed(int nocc, int nvir, const double *oovv, const double *e) {
  double sum = 0;
  const size_t o = nocc, v = nvir;
  for (size_t i = 0; i < o; ++i)
    for (size_t j = 0; j < o; ++j)
      for (size_t a = 0; a < v; ++a)
        for (size_t b = 0; b < v; ++b) {
          double iajb = oovv[((i * o + j) * v + a) * v + b], ibja = oovv[((i * o + j) * v + b) * v + a];
          sum += iajb * (2.0 * iajb - ibja) / (e[i] + e[j] - e[o + a] - e[o + b]);
        }
  return sum;
}

/* computeCiStates, src/configuration_interaction.zig:597-683: transformInts (ao2mo_pp + ao2mo_pppp
 * on the spin-orbital basis), Slater-Condon H_CI over generateDets(nel, nsp, excitations), dsyevd,
 * + V_NN.  Writes the lowest min(nstates, ndet) energies; returns the number of determinants. */
int orc_ci(int n, const double *g, const double *H, const double *C, int nocc, int generalized, const int *exc, int nexc,
           double VN, int nstates, double *energies) {
  int nsp;
  double *g_MS, *C_SO;
  so_transform(n, g, C, generalized, &g_MS, &C_SO, &nsp);
  double *H_SO = malloc(sizeof(double) * nsp * nsp), *H_MS = malloc(sizeof(double) * nsp * nsp);
  if (generalized)
    memcpy(H_SO, H, sizeof(double) * nsp * nsp);
  else
    orc_ao2so_pp(n, H, H_SO);
  orc_ao2mo_pp(nsp, H_SO, C_SO, H_MS);
  const int nel = generalized ? nocc : 2 * nocc;
  Dets d = gen_dets(nel, nsp, exc, nexc);
  const size_t nd = d.ndet;
  double *Hci = malloc(sizeof(double) * nd * nd), *ev = malloc(sizeof(double) * nd), *V = malloc(sizeof(double) * nd * nd);
  for (size_t i = 0; i < nd; ++i)
    for (size_t j = i; j < nd; ++j) {
      double v = slater(d.occ + i * nel, d.occ + j * nel, nel, nsp, H_MS, g_MS);
      Hci[i * nd + j] = v, Hci[j * nd + i] = v;
    }
  orc_eigh((int)nd, Hci, ev, V);
  for (int k = 0; k < nstates && k < (int)nd; ++k) energies[k] = ev[k] + VN;
  free(Hci), free(ev), free(V), free(d.occ), free(H_SO), free(H_MS), free(g_MS), free(C_SO);
  return (int)nd;
}
