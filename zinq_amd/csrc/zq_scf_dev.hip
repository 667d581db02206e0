// Device-resident SCF: the loop of src/hartree_fock.zig:738-875 with every n x n step on the GPU.
//
//   per iteration (all on the context's stream, FP64):
//     F      = H + G(P)                           J/K kernels (getFock, :686-735)
//     E      = 1/2 sum P (H + F) + V_NN           deterministic block reduction (getEnergy, :643-655)
//     err    = F P S - (F P S)^T                   MFMA GEMMs (getError, :658-683)
//     B row  = <err_new, err_i>                   deterministic reductions; the <= 9 x 9 DIIS system
//                                                 is solved on the host (diis, :176-227) and the
//     F_ext  = sum c_i F_i, symmetrised           combination runs on the GPU
//     P      = f X D X^T, D = theta(mu - F')      F' = X^T F_ext X in the orthonormal basis X = L^-T
//                                                 (S = L L^T); D, the projector onto the nocc lowest
//                                                 eigenvectors of F', by SP2 purification (GEMMs only).
//                                                 This is the density of getDensity (:621-640) after the
//                                                 aufbau diagonalisation of geigh; without a HOMO-LUMO
//                                                 gap SP2 does not settle and that iteration falls back
//                                                 to rocSOLVER dsyevd.
//     rms(dP)
//   After the loop one dsyevd of the last diagonalised Fock matrix gives C and e (F C = S C e), and
//   P is rebuilt from C, as the reference returns them.
// Convergence test and DIIS bookkeeping follow the reference exactly (history <= diis, reset on a
// singular DIIS matrix, |dE| and rms(dP) below threshold with i > 0).

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <vector>

#include "../../include/zinq_scf.h"
#include "zq_internal.h"
#include "zq_linalg.h"

namespace {

constexpr int kRedBlocks = 64;

// partial[b] = sum over a fixed strided subset of elements of x[i]*y[i] (or of terms), deterministic.
__global__ __launch_bounds__(256) void k_energy_part(const double *__restrict__ P, int npad, const double *__restrict__ H,
                                                     const double *__restrict__ F, int n, double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t e = blockIdx.x * 256 + threadIdx.x; e < size_t(n) * n; e += size_t(gridDim.x) * 256) {
    const int i = int(e / n), j = int(e % n);
    s += 0.5 * P[size_t(i) * npad + j] * (H[e] + F[e]);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out[k] = sum of the kRedBlocks partials of slot k, in order.
__global__ void k_finish(const double *__restrict__ part, int nslots, double *__restrict__ out) {
  const int k = threadIdx.x;
  if (k < nslots) {
    double s = 0.0;
    for (int b = 0; b < kRedBlocks; ++b) s += part[k * kRedBlocks + b];
    out[k] = s;
  }
}

// err = FPS - (FPS)^T
__global__ void k_antisym(const double *__restrict__ FPS, int n, double *__restrict__ err) {
  for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < size_t(n) * n; e += size_t(gridDim.x) * blockDim.x) {
    const int i = int(e / n), j = int(e % n);
    err[e] = FPS[e] - FPS[size_t(j) * n + i];
  }
}

// part[slot*kRedBlocks + b]: dot(err_new, err_slot) over a strided subset (blockIdx.y = slot).
__global__ __launch_bounds__(256) void k_dots(const double *__restrict__ hist, const double *__restrict__ enew, size_t nn,
                                              double *__restrict__ part) {
  __shared__ double red[256];
  const double *ei = hist + size_t(blockIdx.y) * nn;
  double s = 0.0;
  for (size_t e = blockIdx.x * 256 + threadIdx.x; e < nn; e += size_t(gridDim.x) * 256) s += enew[e] * ei[e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.y * kRedBlocks + blockIdx.x] = red[0];
}

// F = sum_i c_i Fhist[slot_i], then F <- (F + F^T)/2.  One thread per (i <= j) pair... simple form:
// each thread writes F[i][j] and F[j][i] for i <= j from the same sum.
__global__ void k_combine(const double *__restrict__ hist, const int *__restrict__ slots, const double *__restrict__ c, int m,
                          int n, double *__restrict__ F) {
  for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < size_t(n) * n; e += size_t(gridDim.x) * blockDim.x) {
    const int i = int(e / n), j = int(e % n);
    if (j < i) continue;
    double a = 0.0, b = 0.0;
    for (int k = 0; k < m; ++k) {
      const double *Fk = hist + size_t(slots[k]) * n * n;
      a += c[k] * Fk[size_t(i) * n + j];
      b += c[k] * Fk[size_t(j) * n + i];
    }
    const double avg = (a + b) / 2;
    F[size_t(i) * n + j] = avg;
    F[size_t(j) * n + i] = avg;
  }
}

// P_new[i][j] = f sum_{k<o} C[i][k] C[j][k] (padded npad stride); part = sum (P_new - P_old)^2.
__global__ __launch_bounds__(256) void k_density(const double *__restrict__ C, int n, int nocc, double factor,
                                                 double *__restrict__ P, int npad, double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t e = blockIdx.x * 256 + threadIdx.x; e < size_t(n) * n; e += size_t(gridDim.x) * 256) {
    const int i = int(e / n), j = int(e % n);
    double v = 0.0;
    for (int k = 0; k < nocc; ++k) v += factor * C[size_t(i) * n + k] * C[size_t(j) * n + k];
    const double d = v - P[size_t(i) * npad + j];
    s += d * d;
    P[size_t(i) * npad + j] = v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// P[i][j] (padded npad stride) = Pn[i][j]; part = sum (Pn - P_old)^2.
__global__ __launch_bounds__(256) void k_store_density(const double *__restrict__ Pn, int n, double *__restrict__ P, int npad,
                                                       double *__restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (size_t e = blockIdx.x * 256 + threadIdx.x; e < size_t(n) * n; e += size_t(gridDim.x) * 256) {
    const int i = int(e / n), j = int(e % n);
    const double v = Pn[e];
    const double d = v - P[size_t(i) * npad + j];
    s += d * d;
    P[size_t(i) * npad + j] = v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

struct Dev {
  zq_ctx *ctx;
  int n;
  size_t nn;
  double *S, *X, *XT, *C, *T, *Fp, *W, *FP, *FPS, *Pn, *Fh, *Eh, *part, *scal, *coef, *eig, *offd;
  int *slots, *info;
  int sp2_hint = 24;  // SP2 steps of the previous density, the first convergence check comes after it
  int sp2_steps = 0, sp2_calls = 0, sp2_fallbacks = 0;
};

int alloc(zq_ctx *ctx, Dev &d, int n, int maxh) {
  d.ctx = ctx, d.n = n, d.nn = size_t(n) * n;
  const size_t nn = d.nn;
  struct {
    const char *name;
    double **p;
    size_t cnt;
  } bufs[] = {{"scf_S", &d.S, nn},       {"scf_X", &d.X, nn},     {"scf_XT", &d.XT, nn},   {"scf_C", &d.C, nn},
              {"scf_T", &d.T, nn},       {"scf_Fp", &d.Fp, nn},   {"scf_W", &d.W, nn},     {"scf_FP", &d.FP, nn},
              {"scf_FPS", &d.FPS, nn},   {"scf_Pn", &d.Pn, nn},   {"scf_Fh", &d.Fh, nn * maxh},
              {"scf_Eh", &d.Eh, nn * maxh}, {"scf_part", &d.part, size_t(kRedBlocks) * 16},
              {"scf_scal", &d.scal, 16}, {"scf_coef", &d.coef, 16}, {"scf_eig", &d.eig, size_t(n)},
              {"scf_offd", &d.offd, size_t(n)}};
  for (auto &b : bufs) {
    *b.p = (double *)zq_scratch(ctx, b.name, b.cnt * sizeof(double));
    if (!*b.p) return zq_fail(ctx, ZQ_ERR_OOM, std::string("device SCF buffer ") + b.name);
  }
  d.slots = (int *)zq_scratch(ctx, "scf_slots", 64 * sizeof(int));
  d.info = (int *)zq_scratch(ctx, "scf_info", 4 * sizeof(int));
  if (!d.slots || !d.info) return zq_fail(ctx, ZQ_ERR_OOM, "device SCF buffers");
  return ZQ_OK;
}

rocblas_handle solver(zq_ctx *ctx) {
  if (!ctx->solver) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    rocblas_set_stream(h, ctx->stream);
    ctx->solver = h;
  }
  return (rocblas_handle)ctx->solver;
}

// Fp = X^T A X for the symmetric n x n A (X^T stored as XT, so both products are A B^T forms).
int to_orthonormal(Dev &d, const double *A) {
  const int n = d.n;
  if (int rc = zq_gemm_nt(d.ctx, d.XT, n, A, n, d.T, n, n, n, n, 1.0, "scf_gemm")) return rc;  // T = X^T A
  return zq_gemm_nt(d.ctx, d.T, n, d.XT, n, d.Fp, n, n, n, n, 1.0, "scf_gemm");                 // Fp = T X
}

// Eigen-decomposition of Fp (symmetric, column-major == row-major) by rocSOLVER dsyevd, C = X V,
// eigenvalues ascending in d.eig.  Fp is overwritten by the eigenvectors.
int diag_orthonormal(Dev &d, double *Cout) {
  zq_ctx *ctx = d.ctx;
  const int n = d.n;
  rocblas_handle h = solver(ctx);
  if (!h) return zq_fail(ctx, ZQ_ERR_LAPACK, "rocblas handle");
  zq_prof_begin(ctx, "scf_eig");
  const rocblas_status st =
      rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, d.Fp, n, d.eig, d.offd, d.info);
  zq_prof_end(ctx);
  if (st != rocblas_status_success) return zq_fail(ctx, ZQ_ERR_LAPACK, "rocsolver_dsyevd");
  // eigenvector k is row k of the column-major result: C[i][k] = sum_p X[i][p] V[k][p]
  return zq_gemm_nt(ctx, d.X, n, d.Fp, n, Cout, n, n, n, n, 1.0, "scf_gemm");
}

// New density from the symmetric Fock matrix A into ctx->d_P, with part[0..kRedBlocks) = sum dP^2.
int new_density(Dev &d, const double *A, int nocc, double factor) {
  zq_ctx *ctx = d.ctx;
  const int n = d.n, np = ctx->npad;
  if (int rc = to_orthonormal(d, A)) return rc;
  const double *D = nullptr;
  int ld = 0, steps = 0;
  const int rc = zq_sp2_density(ctx, d.Fp, n, nocc, 1e-9, 100, &d.sp2_hint, &D, &ld, &steps);
  ++d.sp2_calls;
  d.sp2_steps += steps;
  if (rc == ZQ_OK) {
    if (int r2 = zq_gemm_nt(ctx, d.X, n, D, ld, d.T, n, n, n, n, 1.0, "scf_gemm")) return r2;        // T = X D
    if (int r2 = zq_gemm_nt(ctx, d.T, n, d.X, n, d.Pn, n, n, n, n, factor, "scf_gemm")) return r2;  // f T X^T
    hipLaunchKernelGGL(k_store_density, dim3(kRedBlocks), dim3(256), 0, ctx->stream, d.Pn, n, ctx->d_P, np, d.part);
    return ZQ_OK;
  }
  if (rc != ZQ_ERR_NOT_CONVERGED) return rc;
  // no gap at the Fermi level within 100 SP2 steps: aufbau occupation of the diagonalised F'
  ++d.sp2_fallbacks;
  if (int r2 = diag_orthonormal(d, d.C)) return r2;
  hipLaunchKernelGGL(k_density, dim3(kRedBlocks), dim3(256), 0, ctx->stream, d.C, n, nocc, factor, ctx->d_P, np, d.part);
  return ZQ_OK;
}

int sync_read(zq_ctx *ctx, const double *dsrc, double *host, size_t cnt) {
  ZQ_TRY(ctx, hipMemcpyAsync(host, dsrc, cnt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ZQ_OK;
}

}  // namespace

void zq_release_solver(zq_ctx *ctx) {
  if (ctx && ctx->solver) {
    rocblas_destroy_handle((rocblas_handle)ctx->solver);
    ctx->solver = nullptr;
  }
}

extern "C" int zq_scf_run_device(zq_ctx *ctx, size_t n_, size_t nocc_, const double *H, const double *S, double VN,
                                 const zq_scf_options *opt, double *C, double *e, double *P, double *F,
                                 zq_scf_result *res) {
  if (!ctx || !H || !S || !opt || !C || !e || !P || !F || !res || n_ == 0 || nocc_ > n_)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_scf_run_device: bad arguments");
  if (!ctx->eri || size_t(ctx->n) != n_) return zq_fail(ctx, ZQ_ERR_STATE, "zq_scf_run_device: ERIs of this size not loaded");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  const int n = int(n_), nocc = int(nocc_), np = ctx->npad;
  const bool generalized = opt->generalized != 0;
  const double x_exch = generalized ? 1.0 : 0.5, factor = generalized ? 1.0 : 2.0;
  const int maxh = opt->diis > 0 ? opt->diis : 1;
  Dev d;
  if (int rc = alloc(ctx, d, n, maxh)) return rc;
  const size_t nn = d.nn;
  const int gblocks = int(std::min<size_t>(1024, (nn + 255) / 256));

  // X = L^-T (S = L L^T), the orthonormal starting basis of the core guess
  std::vector<double> hX(nn), hXT(nn);
  if (zq::chol_inv_t(n, S, hX.data())) return zq_fail(ctx, ZQ_ERR_LAPACK, "overlap matrix not positive definite");
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) hXT[size_t(j) * n + i] = hX[size_t(i) * n + j];
  ZQ_TRY(ctx, hipMemcpyAsync(d.S, S, nn * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(d.X, hX.data(), nn * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(d.XT, hXT.data(), nn * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(ctx->d_H, H, nn * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemsetAsync(ctx->d_P, 0, size_t(np) * np * 8, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));

  // core-Hamiltonian guess, src/hartree_fock.zig:425-431
  if (int rc = new_density(d, ctx->d_H, nocc, factor)) return rc;

  std::deque<int> hist;  // ring-buffer slots of the DIIS history, oldest first
  std::vector<double> Bfull(64 * 64, 0.0);  // B[slot][slot]
  double e_old = VN, e_new = VN;
  const double *Fout = ctx->d_F;  // the Fock matrix diagonalised last (DIIS-extrapolated when DIIS ran)
  int it;
  bool converged = false;
  for (it = 0; it < opt->iterations; ++it) {
    // F = H + G(P); E
    if (int rc = zq_fock_jk_device(ctx, ctx->d_P, ctx->d_H, x_exch, ctx->d_F)) return rc;
    hipLaunchKernelGGL(k_energy_part, dim3(kRedBlocks), dim3(256), 0, ctx->stream, ctx->d_P, np, ctx->d_H, ctx->d_F, n,
                       d.part);
    const double *Fdiag = ctx->d_F;
    double hbuf[16];
    if (opt->diis > 0) {
      // err = FPS - (FPS)^T into the next history slot; F copied alongside
      // F P (P symmetric) and (F P) S (S symmetric) as A B^T products
      if (int rc = zq_gemm_nt(ctx, ctx->d_F, n, ctx->d_P, np, d.FP, n, n, n, n, 1.0, "scf_gemm")) return rc;
      if (int rc = zq_gemm_nt(ctx, d.FP, n, d.S, n, d.FPS, n, n, n, n, 1.0, "scf_gemm")) return rc;
      int slot;
      if (int(hist.size()) >= opt->diis) {
        slot = hist.front();
        hist.pop_front();
      } else {
        slot = 0;
        while (std::find(hist.begin(), hist.end(), slot) != hist.end()) ++slot;
      }
      hipLaunchKernelGGL(k_antisym, dim3(gblocks), dim3(256), 0, ctx->stream, d.FPS, n, d.Eh + size_t(slot) * nn);
      ZQ_TRY(ctx, hipMemcpyAsync(d.Fh + size_t(slot) * nn, ctx->d_F, nn * 8, hipMemcpyDeviceToDevice, ctx->stream));
      hist.push_back(slot);
      const int m = int(hist.size());
      hipLaunchKernelGGL(k_dots, dim3(kRedBlocks, maxh), dim3(256), 0, ctx->stream, d.Eh, d.Eh + size_t(slot) * nn, nn,
                         d.part + kRedBlocks);
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, ctx->stream, d.part, 1 + maxh, d.scal);
      if (int rc = sync_read(ctx, d.scal, hbuf, 1 + maxh)) return rc;
      e_new = hbuf[0] + VN;
      for (int k = 0; k < maxh; ++k) Bfull[slot * 64 + k] = Bfull[k * 64 + slot] = hbuf[1 + k];
      if (m >= 2) {
        const int dd = m + 1;
        std::vector<double> B(size_t(dd) * dd, 0.0), c(dd, 0.0);
        std::vector<int> piv(dd);
        for (int i = 0; i < m; ++i)
          for (int j = 0; j < m; ++j) B[size_t(i) * dd + j] = Bfull[hist[i] * 64 + hist[j]];
        for (int i = 0; i < m; ++i) B[size_t(i) * dd + m] = B[size_t(m) * dd + i] = -1.0;
        c[m] = -1.0;
        if (zq::getrf(dd, B.data(), piv.data()) == 0) {
          zq::getrs(dd, B.data(), piv.data(), c.data());
          std::vector<int> sl(hist.begin(), hist.end());
          ZQ_TRY(ctx, hipMemcpyAsync(d.slots, sl.data(), m * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
          ZQ_TRY(ctx, hipMemcpyAsync(d.coef, c.data(), m * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
          hipLaunchKernelGGL(k_combine, dim3(gblocks), dim3(256), 0, ctx->stream, d.Fh, d.slots, d.coef, m, n, d.W);
          Fdiag = d.W;
          ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));  // slots/coef staged from host vectors
        } else {
          // singular DIIS matrix: keep only the current entry, src/hartree_fock.zig:821-845
          hist.clear();
          hist.push_back(slot);
        }
      }
    } else {
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, ctx->stream, d.part, 1, d.scal);
      if (int rc = sync_read(ctx, d.scal, hbuf, 1)) return rc;
      e_new = hbuf[0] + VN;
    }
    // new density from the (extrapolated) Fock matrix
    Fout = Fdiag;
    if (int rc = new_density(d, Fdiag, nocc, factor)) return rc;
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, ctx->stream, d.part, 1, d.scal);
    if (int rc = sync_read(ctx, d.scal, hbuf, 1)) return rc;
    const double delta_energy = std::fabs(e_new - e_old);
    const double p_rm = std::sqrt(hbuf[0] / double(nn));
    e_old = e_new;
    if (it > 0 && delta_energy < opt->threshold && p_rm < opt->threshold) {
      converged = true;
      ++it;
      break;
    }
  }
  // outputs: C and e from the last diagonalised F, P rebuilt from C, and that F
  if (int rc = to_orthonormal(d, Fout)) return rc;
  if (int rc = diag_orthonormal(d, d.C)) return rc;
  hipLaunchKernelGGL(k_density, dim3(kRedBlocks), dim3(256), 0, ctx->stream, d.C, n, nocc, factor, ctx->d_P, np, d.part);
  std::vector<double> hP(size_t(np) * np);
  ZQ_TRY(ctx, hipMemcpyAsync(C, d.C, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(e, d.eig, size_t(n) * 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(hP.data(), ctx->d_P, hP.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(F, Fout, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < n; ++i) std::memcpy(P + size_t(i) * n, hP.data() + size_t(i) * np, n * sizeof(double));
  res->energy = e_new;
  res->iterations = converged ? it : opt->iterations;
  return converged ? ZQ_OK : zq_fail(ctx, ZQ_ERR_NOT_CONVERGED, "SCF did not converge");
}
