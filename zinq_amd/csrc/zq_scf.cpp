// Host SCF driver: C++ restatement of the reference's scf loop (src/hartree_fock.zig:738-875) with the
// Fock build on the GPU through zq_fock_jk and every other step in FP64 on the host.

#include <cmath>
#include <cstring>
#include <deque>
#include <vector>

#include "../../include/zinq_scf.h"
#include "zq_linalg.h"

namespace {

using Mat = std::vector<double>;

// getDensity, src/hartree_fock.zig:621-640; returns rms(P_new - P_old).
double get_density(int n, const double *C, int nocc, bool generalized, double *P) {
  const double factor = generalized ? 1.0 : 2.0;
  double sum_sq = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int k = 0; k < nocc; ++k) s += factor * C[size_t(i) * n + k] * C[size_t(j) * n + k];
      const double d = s - P[size_t(i) * n + j];
      sum_sq += d * d;
      P[size_t(i) * n + j] = s;
    }
  return std::sqrt(sum_sq / double(size_t(n) * n));
}

// getEnergy, src/hartree_fock.zig:643-655.
double get_energy(int n, const double *H, const double *F, const double *P) {
  double e = 0;
  for (size_t x = 0; x < size_t(n) * n; ++x) e += 0.5 * P[x] * (H[x] + F[x]);
  return e;
}

// getError, src/hartree_fock.zig:658-683: FPS - (FPS)^T.
void get_error(int n, const double *F, const double *P, const double *S, double *err) {
  Mat FP(size_t(n) * n), FPS(size_t(n) * n);
  zq::gemm(n, F, P, FP.data());
  zq::gemm(n, FP.data(), S, FPS.data());
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) err[size_t(i) * n + j] = FPS[size_t(i) * n + j] - FPS[size_t(j) * n + i];
}

// diis, src/hartree_fock.zig:176-227; false if the LU factorisation fails.
bool diis(int n, const std::deque<Mat> &fh, const std::deque<Mat> &eh, double *F) {
  const int m = int(fh.size());
  if (m < 2) return true;
  const int d = m + 1;
  std::vector<double> B(size_t(d) * d, 0.0), c(d, 0.0);
  std::vector<int> piv(d);
  for (int i = 0; i < m; ++i)
    for (int j = i; j < m; ++j) {
      double s = 0;
      for (size_t k = 0; k < size_t(n) * n; ++k) s += eh[i][k] * eh[j][k];
      B[size_t(i) * d + j] = B[size_t(j) * d + i] = s;
    }
  for (int i = 0; i < m; ++i) B[size_t(i) * d + m] = B[size_t(m) * d + i] = -1.0;
  c[m] = -1.0;
  if (zq::getrf(d, B.data(), piv.data())) return false;
  zq::getrs(d, B.data(), piv.data(), c.data());
  std::memset(F, 0, sizeof(double) * n * n);
  for (int i = 0; i < m; ++i)
    for (size_t k = 0; k < size_t(n) * n; ++k) F[k] += c[i] * fh[i][k];
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double avg = (F[size_t(i) * n + j] + F[size_t(j) * n + i]) / 2;
      F[size_t(i) * n + j] = F[size_t(j) * n + i] = avg;
    }
  return true;
}

// geigh: F C = S C e (LAPACKE_dsygvd in the reference, src/linear_algebra.zig:184-200).
int geigh(int n, const double *A, const double *S, double *C, double *e) {
  Mat B(S, S + size_t(n) * n);
  std::memcpy(C, A, sizeof(double) * n * n);
  return zq::sygv(n, C, B.data(), e) ? ZQ_ERR_LAPACK : ZQ_OK;
}

}  // namespace

extern "C" int zq_host_sygv(size_t n, double *A, double *B, double *w) {
  return zq::sygv(int(n), A, B, w) ? ZQ_ERR_LAPACK : ZQ_OK;
}

extern "C" int zq_scf_run(zq_ctx *ctx, size_t n_, size_t nocc_, const double *H, const double *S, double VN,
                          const zq_scf_options *opt, double *C, double *e, double *P, double *F, zq_scf_result *res) {
  if (!ctx || !H || !S || !opt || !C || !e || !P || !F || !res || n_ == 0 || nocc_ > n_) return ZQ_ERR_INVALID;
  const int n = int(n_), nocc = int(nocc_);
  const bool generalized = opt->generalized != 0;
  const double x_exch = generalized ? 1.0 : 0.5;
  const size_t nn = size_t(n) * n;

  // core-Hamiltonian guess, src/hartree_fock.zig:425-431
  std::memset(P, 0, sizeof(double) * nn);
  if (int rc = geigh(n, H, S, C, e)) return rc;
  get_density(n, C, nocc, generalized, P);

  std::deque<Mat> fh, eh;
  double e_old = VN, e_new = VN;
  for (int it = 0; it < opt->iterations; ++it) {
    if (int rc = zq_fock_jk(ctx, H, P, x_exch, F)) return rc;
    e_new = get_energy(n, H, F, P) + VN;
    if (opt->diis > 0) {
      Mat err(nn);
      get_error(n, F, P, S, err.data());
      if (int(fh.size()) >= opt->diis) fh.pop_front(), eh.pop_front();
      fh.emplace_back(F, F + nn);
      eh.push_back(err);
      if (!diis(n, fh, eh, F)) {
        // reset the history to the current entry, src/hartree_fock.zig:821-845
        fh.clear(), eh.clear();
        fh.emplace_back(F, F + nn);
        Mat err2(nn);
        get_error(n, F, P, S, err2.data());
        eh.push_back(err2);
      }
    }
    if (int rc = geigh(n, F, S, C, e)) return rc;
    const double delta_energy = std::fabs(e_new - e_old);
    const double p_rm = get_density(n, C, nocc, generalized, P);
    e_old = e_new;
    if (it > 0 && delta_energy < opt->threshold && p_rm < opt->threshold) {
      res->energy = e_new;
      res->iterations = it + 1;
      return ZQ_OK;
    }
  }
  res->energy = e_new;
  res->iterations = opt->iterations;
  return ZQ_ERR_NOT_CONVERGED;
}
