// Grouped RHF + MP2 of many small molecules (include/zinq_batch.h).
//
// Per SCF iteration, for every molecule that has not converged, three launches captured in one graph:
//   zq_jk_records / zq_jk_reduce over the grouped work plan     F = H + G(P)   (getFock, hartree_fock.zig:686-735)
//   k_batch_step, one workgroup per molecule:
//     E = 1/2 sum P (H + F) + V_NN                                (getEnergy, :643-655)
//     err = F P S - (F P S)^T, history, bordered LU, F_ext        (getError :658-683, diis :176-227, :810-845)
//     F' = X^T F_ext X, SP2 purification -> D, P = 2 X D X^T       (geigh + getDensity, :621-640)
//     rms(dP), |dE|, convergence test                            (:853-870)
//   k_batch_count: molecules done, read by the host one iteration late.
// Every per-molecule matrix lives in one workspace with leading dimension mp = n rounded up to 16 and
// zero padding, so the small GEMMs (zq_tile.h, A B^T on FP64 MFMA, one wave per 16 x 16 tile) never
// need bounds checks.  After convergence: F' of the last extrapolated F for each molecule, rocSOLVER
// dsyevd_strided_batched per distinct n, C = X V, and the MP2 energy from the OVOV transform of
// zq_transform.hip through a context that views the molecule's resident quartets.

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/zinq_batch.h"
#include "zq_internal.h"
#include "zq_linalg.h"
#include "zq_tile.h"

using namespace zqt;

namespace {

constexpr int kMaxN = 96;
constexpr int kHist = 8;
constexpr int kMats = 13;
enum { M_H, M_F, M_P, M_S, M_X, M_XT, M_W, M_T, M_FP, M_FPS, M_FPR, M_XS, M_X2 };
enum { MODE_GUESS = 0, MODE_ITER = 1 };

struct BMol {
  long long ws;  // workspace offset (doubles): kMats + 2 kHist matrices of mp x mp
  int n, mp, nocc, pad;
  double VN;
};

struct BState {
  int done, it, status, hist_n, cur, combine, sp2_steps, pad;
  int hist[kHist];
  double e_old, e_new, rms, dE;
  double coef[kHist + 1];
  double B[kHist * kHist];
};

template <int NT>
__device__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NT / 64; ++k) s += red[k];
  return s;
}

template <int NT>
__device__ void block_minmax(double &lo, double &hi, double *red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = lo, red[NT / 64 + w] = hi;
  __syncthreads();
  lo = red[0], hi = red[NT / 64];
#pragma unroll
  for (int k = 1; k < NT / 64; ++k) lo = fmin(lo, red[k]), hi = fmax(hi, red[NT / 64 + k]);
}

// C = alpha A B^T, all mp x mp with leading dimension mp (mp a multiple of 16); the workgroup's waves
// take the 16 x 16 tiles in turn.  Ends with a barrier.
template <int NT>
__device__ void wg_gemm_nt(const double *A, const double *B, double *C, int mp, double alpha) {
  const int T = mp >> 4, waves = NT / 64, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  for (int t = wave; t < T * T; t += waves) {
    const int ti = t / T, tj = t - ti * T;
    const d4 acc = tile_nt<true>(A, mp, B, mp, ti * 16, tj * 16, mp, mp, mp);
#pragma unroll
    for (int j = 0; j < 4; ++j) C[size_t(ti * 16 + 4 * j + q) * mp + tj * 16 + r] = alpha * acc[j];
  }
  __syncthreads();
}

// The 16 x 16 tile (i0, j0) of X X^T for the symmetric mp x mp X held in LDS (leading dimension ld):
// lane l (r = l % 16, q = l / 16) supplies X[i0 + r][k + q] and X[j0 + r][k + q] to each
// v_mfma_f64_16x16x4f64; accumulator element j is row i0 + 4 j + q, column j0 + r.
__device__ inline d4 lds_tile_sq(const double *Xl, int ld, int mp, int i0, int j0) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const double *a = Xl + (i0 + r) * ld + q, *b = Xl + (j0 + r) * ld + q;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < mp; k += 4) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[k], b[k], acc, 0, 0, 0);
  return acc;
}

// Bordered DIIS system (m + 1) x (m + 1) by LU with partial pivoting, the steps of zq::getrf/getrs.
__device__ bool lu_solve(int n, double *A, double *b) {
  int piv[kHist + 1];
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + k]) > best) best = fabs(A[i * n + k]), p = i;
    piv[k] = p;
    if (best == 0.0) return false;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        const double t = A[k * n + j];
        A[k * n + j] = A[p * n + j], A[p * n + j] = t;
      }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      A[i * n + k] = f;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
    }
  }
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) {
      const double t = b[k];
      b[k] = b[piv[k]], b[piv[k]] = t;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) b[i] -= A[i * n + j] * b[j];
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) b[i] -= A[i * n + j] * b[j];
    b[i] /= A[i * n + i];
  }
  return true;
}

// One SCF step (MODE_ITER) or the core-Hamiltonian guess (MODE_GUESS) of molecule order[blockIdx.x].
// The SP2 iterate X lives in dynamic LDS (mp x (mp + 1) doubles, sized for the batch's largest mp) for
// the whole step loop: each wave computes its 16 x 16 tiles of X^2 from LDS into registers, and after a
// barrier writes its tiles' next X back in place.
template <int NT>
__global__ __launch_bounds__(NT) void k_batch_step(double *__restrict__ ws, const BMol *__restrict__ mols,
                                                   const int *__restrict__ order, BState *__restrict__ st,
                                                   int *__restrict__ done, int mode, int diis, double thr, int maxit) {
  __shared__ double red[2 * (NT / 64)];
  extern __shared__ double Xl[];
  __shared__ double Am[(kHist + 1) * (kHist + 1)], c[kHist + 1];  // the bordered DIIS system (thread 0)
  const int mi = order[blockIdx.x], tid = threadIdx.x;
  if (mode == MODE_ITER && done[mi]) return;
  const BMol M = mols[mi];
  BState &S = st[mi];
  const int n = M.n, mp = M.mp;
  const size_t mm = size_t(mp) * mp;
  double *base = ws + M.ws;
  double *H = base + M_H * mm, *F = base + M_F * mm, *P = base + M_P * mm, *Sm = base + M_S * mm;
  double *X = base + M_X * mm, *XT = base + M_XT * mm, *W = base + M_W * mm, *T = base + M_T * mm;
  double *FP = base + M_FP * mm, *FPS = base + M_FPS * mm, *Fpr = base + M_FPR * mm, *Xs = base + M_XS * mm;
  double *Fh = base + kMats * mm, *Eh = Fh + kHist * mm;
  const double *A = H;

  if (mode == MODE_GUESS) {
    if (tid == 0) {
      S.done = 0, S.it = 0, S.status = 0, S.hist_n = 0, S.cur = 0, S.combine = 0, S.sp2_steps = 0;
      S.e_old = M.VN, S.e_new = M.VN, S.rms = 0.0, S.dE = 0.0;
      done[mi] = 0;
    }
  } else {
    // energy of P and F (F = H + G(P) from the grouped J/K)
    double e = 0.0;
    for (int x = tid; x < n * n; x += NT) {
      const int i = x / n, j = x - i * n;
      const size_t o = size_t(i) * mp + j;
      e += 0.5 * P[o] * (H[o] + F[o]);
    }
    e = block_sum<NT>(e, red) + M.VN;
    if (diis > 0) {
      if (tid == 0) {
        int slot;
        if (S.hist_n >= diis) {
          slot = S.hist[0];
          for (int k = 1; k < S.hist_n; ++k) S.hist[k - 1] = S.hist[k];
          --S.hist_n;
        } else {
          slot = 0;
          for (bool used = true; used;) {
            used = false;
            for (int k = 0; k < S.hist_n; ++k) used |= S.hist[k] == slot;
            if (used) ++slot;
          }
        }
        S.hist[S.hist_n++] = slot;
        S.cur = slot;
      }
      wg_gemm_nt<NT>(F, P, FP, mp, 1.0);    // F P   (P symmetric)
      wg_gemm_nt<NT>(FP, Sm, FPS, mp, 1.0);  // F P S (S symmetric)
      const int cur = S.cur, hn = S.hist_n;
      double *Ec = Eh + cur * mm, *Fc = Fh + cur * mm;
      for (size_t x = tid; x < mm; x += NT) {
        const int i = int(x / mp), j = int(x - size_t(i) * mp);
        Ec[x] = FPS[x] - FPS[size_t(j) * mp + i];
        Fc[x] = F[x];
      }
      __syncthreads();
      for (int k = 0; k < hn; ++k) {
        const double *Ek = Eh + S.hist[k] * mm;
        double d = 0.0;
        for (size_t x = tid; x < mm; x += NT) d += Ec[x] * Ek[x];
        d = block_sum<NT>(d, red);
        if (tid == 0) S.B[cur * kHist + S.hist[k]] = S.B[S.hist[k] * kHist + cur] = d;
      }
      __syncthreads();
      if (tid == 0) {
        S.combine = 0;
        if (hn >= 2) {
          const int dd = hn + 1;
          for (int i = 0; i < dd * dd; ++i) Am[i] = 0.0;
          for (int i = 0; i < dd; ++i) c[i] = 0.0;
          for (int i = 0; i < hn; ++i)
            for (int j = 0; j < hn; ++j) Am[i * dd + j] = S.B[S.hist[i] * kHist + S.hist[j]];
          for (int i = 0; i < hn; ++i) Am[i * dd + hn] = Am[hn * dd + i] = -1.0;
          c[hn] = -1.0;
          if (lu_solve(dd, Am, c)) {
            for (int i = 0; i < hn; ++i) S.coef[i] = c[i];
            S.combine = 1;
          } else {
            S.hist[0] = cur;
            S.hist_n = 1;
          }
        }
      }
      __syncthreads();
    }
    const int comb = diis > 0 ? S.combine : 0, hn = S.hist_n;
    for (size_t x = tid; x < mm; x += NT) {
      const int i = int(x / mp), j = int(x - size_t(i) * mp);
      if (!comb) {
        W[x] = F[x];
      } else if (j >= i) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < hn; ++k) {
          const double *Fk = Fh + S.hist[k] * mm;
          a += S.coef[k] * Fk[size_t(i) * mp + j];
          b += S.coef[k] * Fk[size_t(j) * mp + i];
        }
        const double avg = (a + b) / 2;
        W[size_t(i) * mp + j] = avg;
        W[size_t(j) * mp + i] = avg;
      }
    }
    if (tid == 0) S.e_new = e;
    __syncthreads();
    A = W;
  }

  // F' = X^T A X
  wg_gemm_nt<NT>(XT, A, T, mp, 1.0);
  wg_gemm_nt<NT>(T, XT, Fpr, mp, 1.0);
  // SP2 on F' (Gershgorin bounds, trace-correcting steps; see zq_dense.hip)
  double lo = 1e300, hi = -1e300;
  for (int i = tid; i < n; i += NT) {
    double rs = 0.0;
    for (int j = 0; j < n; ++j) rs += fabs(Fpr[size_t(j) * mp + i]);
    const double dgl = Fpr[size_t(i) * mp + i];
    rs -= fabs(dgl);
    lo = fmin(lo, dgl - rs), hi = fmax(hi, dgl + rs);
  }
  block_minmax<NT>(lo, hi, red);
  const double pad = 1e-12 * fmax(fabs(lo), fabs(hi));
  lo -= pad, hi += pad;
  const double inv = 1.0 / fmax(hi - lo, 1e-10 * fmax(1.0, fabs(hi)));
  const int ld = mp + 1;  // odd: the 16 rows of a tile fall in distinct LDS banks
  double t1 = 0.0, t2 = 0.0;
  for (size_t x = tid; x < mm; x += NT) {
    const int i = int(x / mp), j = int(x - size_t(i) * mp);
    double v = 0.0;
    if (i < n && j < n) v = ((i == j ? hi : 0.0) - 0.5 * (Fpr[x] + Fpr[size_t(j) * mp + i])) * inv;
    Xl[i * ld + j] = v;
    t1 += (i == j) ? v : 0.0;
    t2 += v * v;
  }
  t1 = block_sum<NT>(t1, red);
  t2 = block_sum<NT>(t2, red);
  const double N = double(M.nocc);
  constexpr int NW = NT / 64, kTiles = (kMaxN / 16) * (kMaxN / 16), kPer = (kTiles + NW - 1) / NW;
  const int Tt = mp >> 4, wave = tid >> 6, lane = tid & 63, r16 = lane & 15, q16 = lane >> 4;
  int count = -1, steps = 0;
  for (; steps < 100; ++steps) {
    if (count >= 2) break;
    if (count < 0 && t1 - t2 < 1e-9) count = 0;
    if (count >= 0) ++count;
    const bool sq = fabs(t2 - N) <= fabs(2.0 * t1 - t2 - N);
    d4 acc[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int t = wave + u * NW;
      if (t < Tt * Tt) acc[u] = lds_tile_sq(Xl, ld, mp, (t / Tt) * 16, (t % Tt) * 16);
    }
    __syncthreads();  // every tile of X^2 is formed before X is overwritten
    double a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int t = wave + u * NW;
      if (t < Tt * Tt) {
        const int ti = t / Tt, tj = t % Tt;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int i = ti * 16 + 4 * j + q16, c = tj * 16 + r16;
          const double x2 = acc[u][j];
          const double v = sq ? x2 : 2.0 * Xl[i * ld + c] - x2;
          Xl[i * ld + c] = v;
          a1 += (i == c) ? v : 0.0;
          a2 += v * v;
        }
      }
    }
    t1 = block_sum<NT>(a1, red);  // its barriers also order the writes before the next step's reads
    t2 = block_sum<NT>(a2, red);
  }
  if (count < 2) {  // no gap at the Fermi level: this molecule stops here
    if (tid == 0) S.status = 3, S.done = 1, done[mi] = 1;
    return;
  }
  for (size_t x = tid; x < mm; x += NT) {
    const int i = int(x / mp), j = int(x - size_t(i) * mp);
    Xs[x] = Xl[i * ld + j];
  }
  __syncthreads();
  // P = 2 X D X^T, rms(dP)
  wg_gemm_nt<NT>(X, Xs, T, mp, 1.0);
  wg_gemm_nt<NT>(T, X, FP, mp, 2.0);
  double s = 0.0;
  for (size_t x = tid; x < mm; x += NT) {
    const double dlt = FP[x] - P[x];
    s += dlt * dlt;
    P[x] = FP[x];
  }
  s = block_sum<NT>(s, red);
  if (tid == 0) {
    S.sp2_steps = steps;
    if (mode == MODE_ITER) {
      const double rms = sqrt(s / (double(n) * n));
      const double dE = fabs(S.e_new - S.e_old);
      S.e_old = S.e_new;
      S.rms = rms, S.dE = dE;
      const bool conv = S.it > 0 && dE < thr && rms < thr;
      ++S.it;
      if (conv) {
        S.done = 1, S.status = 1;
      } else if (S.it >= maxit) {
        S.done = 1, S.status = 2;
      }
      done[mi] = S.done;
    }
  }
}

__global__ void k_batch_count(const int *__restrict__ done, int m, int *__restrict__ out) {
  __shared__ int red[256];
  int c = 0;
  for (int i = threadIdx.x; i < m; i += 256) c += done[i] != 0;
  red[threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// F' = X^T W X of every molecule into eig (n x n, leading dimension n, at eoff[m]).
template <int NT>
__global__ __launch_bounds__(NT) void k_batch_fprime(double *__restrict__ ws, const BMol *__restrict__ mols,
                                                     double *__restrict__ eig, const long long *__restrict__ eoff) {
  const BMol M = mols[blockIdx.x];
  const int n = M.n, mp = M.mp;
  const size_t mm = size_t(mp) * mp;
  double *base = ws + M.ws;
  double *XT = base + M_XT * mm, *W = base + M_W * mm, *T = base + M_T * mm, *Fpr = base + M_FPR * mm;
  wg_gemm_nt<NT>(XT, W, T, mp, 1.0);
  wg_gemm_nt<NT>(T, XT, Fpr, mp, 1.0);
  double *out = eig + eoff[blockIdx.x];
  for (int x = threadIdx.x; x < n * n; x += NT) {
    const int i = x / n, j = x - i * n;
    out[x] = 0.5 * (Fpr[size_t(i) * mp + j] + Fpr[size_t(j) * mp + i]);
  }
}

// C = X V with the eigenvectors of eig (column-major: eigenvector k at eig[eoff + k*n ...]) into
// cout (n x n, leading dimension n, at coff[m]).
template <int NT>
__global__ __launch_bounds__(NT) void k_batch_coef(double *__restrict__ ws, const BMol *__restrict__ mols,
                                                   const double *__restrict__ eig, const long long *__restrict__ eoff,
                                                   double *__restrict__ cout, const long long *__restrict__ coff) {
  const BMol M = mols[blockIdx.x];
  const int n = M.n, mp = M.mp;
  const size_t mm = size_t(mp) * mp;
  double *base = ws + M.ws;
  double *X = base + M_X * mm, *T = base + M_T * mm, *FP = base + M_FP * mm;
  const double *V = eig + eoff[blockIdx.x];
  for (size_t x = threadIdx.x; x < mm; x += NT) {
    const int k = int(x / mp), p = int(x - size_t(k) * mp);
    T[x] = (k < n && p < n) ? V[size_t(k) * n + p] : 0.0;
  }
  __syncthreads();
  wg_gemm_nt<NT>(X, T, FP, mp, 1.0);
  double *out = cout + coff[blockIdx.x];
  for (int x = threadIdx.x; x < n * n; x += NT) {
    const int i = x / n, j = x - i * n;
    out[x] = FP[size_t(i) * mp + j];
  }
}

struct HostMol {
  int n, nocc;
  double VN;
  std::vector<double> H, S, eri;
};

}  // namespace

struct zq_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<HostMol> mols;
  bool uploaded = false;
  int chunk = 0, nt = 256;
  // device residency
  double *eri = nullptr, *ws = nullptr, *rec = nullptr;
  BMol *d_mols = nullptr;
  int *d_order = nullptr;  // molecules by descending padded size: the step kernel's workgroup order
  BState *d_state = nullptr;
  int *d_done = nullptr, *d_count = nullptr, *d_ptr = nullptr, *d_idx = nullptr;
  JKItem *d_items = nullptr;
  JKMol *d_jkm = nullptr;
  int2 *d_blk = nullptr;
  int nitems = 0, nblk = 0;
  std::vector<long long> eri_off, ws_off;
  size_t eri_count = 0;  // doubles in eri
  int mp_max = 16;       // largest padded size of the batch (the SP2 iterate's LDS)
  int *h_count = nullptr;
  hipGraphExec_t graph = nullptr;
  std::string graph_key;
  rocblas_handle solver = nullptr;
  // the grouped MP2's FP32 copy of the quartets (FP32 forms; made once per upload) and workspace
  float *eri32 = nullptr;
  void *mp2_ws = nullptr;
  size_t mp2_ws_bytes = 0;
  // the eigenpair stage's device buffers (grow-only, kept across runs) and pinned host copies of C, e and
  // the solver's info codes, which a second stream brings back while the MP2 runs
  double *d_eig = nullptr, *d_c = nullptr, *d_w = nullptr, *d_e = nullptr;
  long long *d_eoff = nullptr;
  int *d_info = nullptr;
  size_t eig_cap = 0, w_cap = 0, m_cap = 0;  // doubles of d_eig / d_c, of d_w / d_e; molecules
  double *h_c = nullptr, *h_w = nullptr;
  int *h_info = nullptr;
  size_t hc_cap = 0, hw_cap = 0, hi_cap = 0;
  hipStream_t copy_stream = nullptr;
  hipEvent_t coef_done = nullptr;
  // the eigenpair stage's extra lanes: size groups of the batched eigensolver run side by side, each
  // lane a stream with its own rocSOLVER handle (its own workspace); fork/join events
  static constexpr int kEigLanes = 2;
  hipStream_t eig_lane[kEigLanes] = {};
  rocblas_handle eig_solver[kEigLanes] = {};
  hipEvent_t eig_ev[kEigLanes + 1] = {};
  // results of the last run (C and e of molecule m at c_off[m], w_off[m] of h_c, h_w), and its stage
  // times (host wall clock, ms): SCF, eigenpairs, MP2, total
  std::vector<long long> c_off, w_off;
  std::vector<int> res_n;
  double stage_ms[4] = {0, 0, 0, 0};
};

namespace {

int bfail(zq_batch *b, int code, const std::string &msg) {
  if (b) b->err = msg;
  return code;
}

#define BTRY(b, call)                                                                     \
  do {                                                                                    \
    hipError_t _e = (call);                                                               \
    if (_e != hipSuccess) return bfail(b, ZQ_ERR_LAUNCH, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

void free_plan(zq_batch *b) {
  for (void *p : {(void *)b->rec, (void *)b->d_items, (void *)b->d_ptr, (void *)b->d_idx, (void *)b->d_blk})
    if (p) (void)hipFree(p);
  b->rec = nullptr, b->d_items = nullptr, b->d_ptr = nullptr, b->d_idx = nullptr, b->d_blk = nullptr;
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
}

void free_eig(zq_batch *b) {
  if (b->copy_stream) (void)hipStreamSynchronize(b->copy_stream);  // copies into the pinned buffers
  for (void *p : {(void *)b->d_eig, (void *)b->d_c, (void *)b->d_w, (void *)b->d_e, (void *)b->d_eoff, (void *)b->d_info})
    if (p) (void)hipFree(p);
  b->d_eig = b->d_c = b->d_w = b->d_e = nullptr;
  b->d_eoff = nullptr, b->d_info = nullptr;
  b->eig_cap = b->w_cap = b->m_cap = 0;
  for (void *p : {(void *)b->h_c, (void *)b->h_w, (void *)b->h_info})
    if (p) (void)hipHostFree(p);
  b->h_c = b->h_w = nullptr, b->h_info = nullptr;
  b->hc_cap = b->hw_cap = b->hi_cap = 0;
}

// Grow-only device and pinned host buffers of the eigenpair stage.
int eig_buffers(zq_batch *b, size_t etot, size_t wtot, size_t M) {
  if (etot > b->eig_cap) {
    for (double **p : {&b->d_eig, &b->d_c}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    if (b->h_c) (void)hipHostFree(b->h_c);
    b->h_c = nullptr, b->eig_cap = b->hc_cap = 0;
    BTRY(b, hipMalloc(&b->d_eig, etot * sizeof(double)));
    BTRY(b, hipMalloc(&b->d_c, etot * sizeof(double)));
    BTRY(b, hipHostMalloc((void **)&b->h_c, etot * sizeof(double), hipHostMallocDefault));
    b->eig_cap = b->hc_cap = etot;
  }
  if (wtot > b->w_cap) {
    for (double **p : {&b->d_w, &b->d_e}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    if (b->h_w) (void)hipHostFree(b->h_w);
    b->h_w = nullptr, b->w_cap = b->hw_cap = 0;
    BTRY(b, hipMalloc(&b->d_w, wtot * sizeof(double)));
    BTRY(b, hipMalloc(&b->d_e, wtot * sizeof(double)));
    BTRY(b, hipHostMalloc((void **)&b->h_w, wtot * sizeof(double), hipHostMallocDefault));
    b->w_cap = b->hw_cap = wtot;
  }
  if (M > b->m_cap) {
    if (b->d_eoff) (void)hipFree(b->d_eoff);
    if (b->d_info) (void)hipFree(b->d_info);
    if (b->h_info) (void)hipHostFree(b->h_info);
    b->d_eoff = nullptr, b->d_info = nullptr, b->h_info = nullptr, b->m_cap = b->hi_cap = 0;
    BTRY(b, hipMalloc(&b->d_eoff, M * sizeof(long long)));
    BTRY(b, hipMalloc(&b->d_info, M * sizeof(int)));
    BTRY(b, hipHostMalloc((void **)&b->h_info, M * sizeof(int), hipHostMallocDefault));
    b->m_cap = b->hi_cap = M;
  }
  if (!b->copy_stream) BTRY(b, hipStreamCreateWithFlags(&b->copy_stream, hipStreamNonBlocking));
  if (!b->coef_done) BTRY(b, hipEventCreateWithFlags(&b->coef_done, hipEventDisableTiming));
  return ZQ_OK;
}

void free_all(zq_batch *b) {
  free_plan(b);
  free_eig(b);
  for (void *p : {(void *)b->eri, (void *)b->ws, (void *)b->d_mols, (void *)b->d_order, (void *)b->d_state,
                  (void *)b->d_done, (void *)b->d_count, (void *)b->d_jkm, (void *)b->eri32, b->mp2_ws})
    if (p) (void)hipFree(p);
  b->eri = nullptr, b->ws = nullptr, b->d_mols = nullptr, b->d_order = nullptr, b->d_state = nullptr;
  b->d_done = nullptr;
  b->d_count = nullptr, b->d_jkm = nullptr, b->eri32 = nullptr, b->mp2_ws = nullptr, b->mp2_ws_bytes = 0;
  b->uploaded = false;
}

int mp_of(int n) { return (n + 15) / 16 * 16; }
int nb_of(int n) { return (n + ZQ_BLK - 1) / ZQ_BLK; }

// The grouped J/K plan for item size `chunk`: items, records and CSR rows of all molecules concatenated.
int build_plan(zq_batch *b, int chunk) {
  free_plan(b);
  std::vector<JKItem> items;
  std::vector<int> ptr(1, 0), idx;
  std::vector<int2> blk;
  size_t rec_base = 0;
  for (size_t m = 0; m < b->mols.size(); ++m) {
    const int nb = nb_of(b->mols[m].n), Nb = nb * (nb + 1) / 2;
    const size_t nq = size_t(Nb) * (Nb + 1) / 2;
    JKPlan pl;
    zq_jk_plan(Nb, nq, chunk, pl);
    for (JKItem it : pl.items) {
      it.rec0 += int(rec_base);
      it.mol = int(m);
      items.push_back(it);
    }
    for (int o = 0; o < Nb; ++o) {
      for (int e = pl.ptr[o]; e < pl.ptr[o + 1]; ++e) {
        const int code = pl.idx[e];
        idx.push_back(int((size_t(code >> 1) + rec_base) << 1) | (code & 1));
      }
      ptr.push_back(int(idx.size()));
      blk.push_back(make_int2(int(m), o));
    }
    rec_base += pl.nrecords;
    if (rec_base >= (size_t(1) << 30)) return bfail(b, ZQ_ERR_INVALID, "batch J/K plan: too many records");
  }
  b->nitems = int(items.size());
  b->nblk = int(blk.size());
  BTRY(b, hipMalloc(&b->rec, std::max<size_t>(1, rec_base) * 64 * sizeof(double)));
  BTRY(b, hipMalloc(&b->d_items, items.size() * sizeof(JKItem)));
  BTRY(b, hipMalloc(&b->d_ptr, ptr.size() * sizeof(int)));
  BTRY(b, hipMalloc(&b->d_idx, std::max<size_t>(1, idx.size()) * sizeof(int)));
  BTRY(b, hipMalloc(&b->d_blk, blk.size() * sizeof(int2)));
  // the batch stream is non-blocking: uploads go on it, so its kernels see them
  BTRY(b, hipMemcpyAsync(b->d_items, items.data(), items.size() * sizeof(JKItem), hipMemcpyHostToDevice, b->stream));
  BTRY(b, hipMemcpyAsync(b->d_ptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
  if (!idx.empty())
    BTRY(b, hipMemcpyAsync(b->d_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
  BTRY(b, hipMemcpyAsync(b->d_blk, blk.data(), blk.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream));
  BTRY(b, hipStreamSynchronize(b->stream));  // the plan's host vectors go out of scope
  b->chunk = chunk;
  return ZQ_OK;
}

// ERIs into quartets, H/S/X/X^T into the workspace, the per-molecule tables, the plan.
int upload(zq_batch *b) {
  if (b->uploaded) return ZQ_OK;
  if (b->mols.empty()) return bfail(b, ZQ_ERR_STATE, "empty batch");
  BTRY(b, hipSetDevice(b->device));
  const size_t M = b->mols.size();
  b->eri_off.assign(M, 0);
  b->ws_off.assign(M, 0);
  size_t eri_total = 0, ws_total = 0, canon_max = 0;
  std::vector<BMol> bm(M);
  std::vector<JKMol> jm(M);
  for (size_t m = 0; m < M; ++m) {
    const HostMol &h = b->mols[m];
    const int nb = nb_of(h.n), Nb = nb * (nb + 1) / 2, mp = mp_of(h.n);
    const size_t nq = size_t(Nb) * (Nb + 1) / 2, mm = size_t(mp) * mp;
    b->eri_off[m] = (long long)eri_total;
    b->ws_off[m] = (long long)ws_total;
    eri_total += nq * ZQ_QUARTET;
    ws_total += (kMats + 2 * kHist) * mm;
    canon_max = std::max(canon_max, h.eri.size());
    bm[m] = BMol{(long long)b->ws_off[m], h.n, mp, h.nocc, 0, h.VN};
    b->mp_max = std::max(b->mp_max, mp);
    jm[m] = JKMol{b->eri_off[m], b->ws_off[m] + M_P * (long long)mm, b->ws_off[m] + M_H * (long long)mm,
                  b->ws_off[m] + M_F * (long long)mm, mp, h.n};
  }
  BTRY(b, hipMalloc(&b->eri, eri_total * sizeof(double)));
  b->eri_count = eri_total;
  BTRY(b, hipMalloc(&b->ws, ws_total * sizeof(double)));
  BTRY(b, hipMalloc(&b->d_mols, M * sizeof(BMol)));
  BTRY(b, hipMalloc(&b->d_order, M * sizeof(int)));
  BTRY(b, hipMalloc(&b->d_state, M * sizeof(BState)));
  BTRY(b, hipMalloc(&b->d_done, M * sizeof(int)));
  BTRY(b, hipMalloc(&b->d_count, sizeof(int)));
  BTRY(b, hipMalloc(&b->d_jkm, M * sizeof(JKMol)));
  BTRY(b, hipMemsetAsync(b->ws, 0, ws_total * sizeof(double), b->stream));
  BTRY(b, hipMemsetAsync(b->d_done, 0, M * sizeof(int), b->stream));
  BTRY(b, hipMemcpyAsync(b->d_mols, bm.data(), M * sizeof(BMol), hipMemcpyHostToDevice, b->stream));
  BTRY(b, hipMemcpyAsync(b->d_jkm, jm.data(), M * sizeof(JKMol), hipMemcpyHostToDevice, b->stream));
  // largest first: a step of a large molecule takes a compute unit for longest, so the small ones fill
  // in behind it rather than leaving a large one for the end (one workgroup per compute unit at 512
  // threads)
  std::vector<int> order(M);
  for (size_t m = 0; m < M; ++m) order[m] = int(m);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return bm[x].mp > bm[y].mp; });
  BTRY(b, hipMemcpyAsync(b->d_order, order.data(), M * sizeof(int), hipMemcpyHostToDevice, b->stream));
  BTRY(b, hipStreamSynchronize(b->stream));
  double *canon = nullptr;
  BTRY(b, hipMalloc(&canon, canon_max * sizeof(double)));
  for (size_t m = 0; m < M; ++m) {
    const HostMol &h = b->mols[m];
    const int nb = nb_of(h.n), Nb = nb * (nb + 1) / 2, mp = mp_of(h.n);
    const size_t nq = size_t(Nb) * (Nb + 1) / 2, mm = size_t(mp) * mp;
    hipError_t e = hipMemcpyAsync(canon, h.eri.data(), h.eri.size() * sizeof(double), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(zq_pack_quartets, dim3(unsigned(nq)), dim3(256), 0, b->stream, canon, b->eri + b->eri_off[m], h.n,
                         nq);
      e = hipStreamSynchronize(b->stream);
    }
    // H, S, X = L^-T, X^T, zero padded to mp
    std::vector<double> stage(4 * mm, 0.0), hX(size_t(h.n) * h.n);
    if (e == hipSuccess && zq::chol_inv_t(h.n, h.S.data(), hX.data())) {
      (void)hipFree(canon);
      return bfail(b, ZQ_ERR_LAPACK, "batch molecule " + std::to_string(m) + ": overlap not positive definite");
    }
    for (int i = 0; i < h.n; ++i)
      for (int j = 0; j < h.n; ++j) {
        stage[0 * mm + size_t(i) * mp + j] = h.H[size_t(i) * h.n + j];
        stage[1 * mm + size_t(i) * mp + j] = h.S[size_t(i) * h.n + j];
        stage[2 * mm + size_t(i) * mp + j] = hX[size_t(i) * h.n + j];
        stage[3 * mm + size_t(j) * mp + i] = hX[size_t(i) * h.n + j];
      }
    double *w = b->ws + b->ws_off[m];
    if (e == hipSuccess) e = hipMemcpyAsync(w + M_H * mm, stage.data(), mm * 8, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(w + M_S * mm, stage.data() + mm, mm * 8, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(w + M_X * mm, stage.data() + 2 * mm, mm * 8, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(w + M_XT * mm, stage.data() + 3 * mm, mm * 8, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);  // stage goes out of scope
    if (e != hipSuccess) {
      (void)hipFree(canon);
      return bfail(b, ZQ_ERR_LAUNCH, std::string("batch upload: ") + hipGetErrorString(e));
    }
  }
  (void)hipFree(canon);
  if (int rc = build_plan(b, b->chunk)) return rc;
  if (!b->h_count) BTRY(b, hipHostMalloc((void **)&b->h_count, 2 * sizeof(int), hipHostMallocDefault));
  b->uploaded = true;
  return ZQ_OK;
}

void launch_step(zq_batch *b, int mode, int diis, double thr, int maxit) {
  const dim3 grid(unsigned(b->mols.size()));
  const size_t lds = size_t(b->mp_max) * (b->mp_max + 1) * sizeof(double);  // the SP2 iterate
  if (b->nt == 512)
    hipLaunchKernelGGL(k_batch_step<512>, grid, dim3(512), lds, b->stream, b->ws, b->d_mols, b->d_order, b->d_state,
                       b->d_done, mode, diis, thr, maxit);
  else
    hipLaunchKernelGGL(k_batch_step<256>, grid, dim3(256), lds, b->stream, b->ws, b->d_mols, b->d_order, b->d_state,
                       b->d_done, mode, diis, thr, maxit);
}

int enqueue_iteration(zq_batch *b, int diis, double thr, int maxit) {
  BTRY(b, zq_jk_group(b->stream, b->eri, b->d_items, b->nitems, b->ws, 0.5, b->rec, b->d_jkm, b->d_done, b->d_ptr,
                      b->d_idx, b->d_blk, b->nblk, b->ws, b->ws));
  launch_step(b, MODE_ITER, diis, thr, maxit);
  hipLaunchKernelGGL(k_batch_count, dim3(1), dim3(256), 0, b->stream, b->d_done, int(b->mols.size()), b->d_count);
  BTRY(b, hipGetLastError());
  return ZQ_OK;
}

// A view context on molecule m's resident quartets, for the transform code of zq_transform.hip.
void bind_view(zq_ctx *v, zq_batch *b, int m) {
  const int n = b->mols[m].n;
  if (v->eri32 && v->eri32_owned) (void)hipFree(v->eri32);  // an FP32 copy of the previous molecule
  v->eri32 = nullptr;
  v->eri32_owned = true;
  v->eri_owned = false;
  v->n = n;
  v->nb = nb_of(n);
  v->npad = v->nb * ZQ_BLK;
  v->Nb = v->nb * (v->nb + 1) / 2;
  v->nquartets = size_t(v->Nb) * (v->Nb + 1) / 2;
  v->eri = b->eri + b->eri_off[m];
}

}  // namespace

extern "C" {

zq_batch *zq_batch_create(int device) {
  ZqGate _gate;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  auto *b = new zq_batch;
  b->device = device;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return nullptr;
  }
  return b;
}

void zq_batch_destroy(zq_batch *b) {
  ZqGate _gate;
  if (!b) return;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  if (b->copy_stream) (void)hipStreamSynchronize(b->copy_stream);
  free_all(b);
  if (b->h_count) (void)hipHostFree(b->h_count);
  if (b->solver) rocblas_destroy_handle(b->solver);
  for (int l = 0; l < zq_batch::kEigLanes; ++l) {
    if (b->eig_lane[l]) (void)hipStreamSynchronize(b->eig_lane[l]);
    if (b->eig_solver[l]) rocblas_destroy_handle(b->eig_solver[l]);
    if (b->eig_lane[l]) (void)hipStreamDestroy(b->eig_lane[l]);
  }
  for (hipEvent_t e : b->eig_ev)
    if (e) (void)hipEventDestroy(e);
  if (b->coef_done) (void)hipEventDestroy(b->coef_done);
  if (b->copy_stream) (void)hipStreamDestroy(b->copy_stream);
  (void)hipStreamDestroy(b->stream);
  delete b;
}

const char *zq_batch_last_error(const zq_batch *b) { return b ? b->err.c_str() : "null batch"; }

int zq_batch_count(const zq_batch *b) { return b ? int(b->mols.size()) : 0; }

int zq_batch_add(zq_batch *b, size_t n, size_t nocc, const double *H, const double *S, double VN,
                 const double *eri_packed) {
  if (!b || !H || !S || !eri_packed) return bfail(b, ZQ_ERR_INVALID, "zq_batch_add: null argument");
  if (n == 0 || n > size_t(kMaxN) || nocc == 0 || nocc >= n)
    return bfail(b, ZQ_ERR_INVALID, "zq_batch_add: need 0 < nocc < n <= 96");
  HostMol h;
  h.n = int(n), h.nocc = int(nocc), h.VN = VN;
  h.H.assign(H, H + n * n);
  h.S.assign(S, S + n * n);
  const size_t np = n * (n + 1) / 2, nu = np * (np + 1) / 2;
  h.eri.assign(eri_packed, eri_packed + nu);
  if (b->uploaded) {
    ZqGate _gate;
    free_all(b);  // residency is rebuilt on the next run
  }
  b->mols.push_back(std::move(h));
  return int(b->mols.size()) - 1;
}

int zq_batch_autotune(zq_batch *b, int chosen[2]) {
  ZqGate _gate;
  if (!b) return ZQ_ERR_INVALID;
  BTRY(b, hipSetDevice(b->device));
  if (int rc = upload(b)) return rc;
  hipEvent_t e0, e1;
  BTRY(b, hipEventCreate(&e0));
  BTRY(b, hipEventCreate(&e1));
  auto time_it = [&](auto &&fn) -> float {
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      (void)hipEventRecord(e0, b->stream);
      fn();
      (void)hipEventRecord(e1, b->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      best = std::min(best, ms);
    }
    return best;
  };
  int best_chunk = 1;
  float best = 1e30f;
  BTRY(b, hipMemsetAsync(b->d_done, 0, b->mols.size() * sizeof(int), b->stream));
  for (int ch : {1, 2, 4, 8, 16}) {
    if (int rc = build_plan(b, ch)) return rc;
    const float ms = time_it([&] {
      (void)zq_jk_group(b->stream, b->eri, b->d_items, b->nitems, b->ws, 0.5, b->rec, b->d_jkm, b->d_done, b->d_ptr,
                        b->d_idx, b->d_blk, b->nblk, b->ws, b->ws);
    });
    if (ms < best) best = ms, best_chunk = ch;
  }
  int best_nt = 256;
  best = 1e30f;
  for (int nt : {256, 512}) {
    b->nt = nt;
    const float ms = time_it([&] { launch_step(b, MODE_GUESS, 8, 1e-8, 100); });
    if (ms < best) best = ms, best_nt = nt;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->nt = best_nt;
  if (int rc = build_plan(b, best_chunk)) return rc;
  BTRY(b, hipStreamSynchronize(b->stream));
  if (chosen) chosen[0] = best_chunk, chosen[1] = best_nt;
  return ZQ_OK;
}

int zq_batch_run(zq_batch *b, const zq_scf_options *opt, int mp2_precision, double *E_rhf, double *E_corr,
                 int *iterations, int *status) {
  ZqGate _gate;
  if (!b || !opt || !E_rhf || !E_corr || !iterations || !status) return bfail(b, ZQ_ERR_INVALID, "zq_batch_run: null argument");
  if (opt->generalized || opt->fock_precision != ZQ_PREC_FP64 || opt->diis < 0 || opt->diis > kHist || opt->iterations < 1)
    return bfail(b, ZQ_ERR_INVALID, "zq_batch_run: RHF, FP64 Fock builds, diis <= 8, iterations >= 1");
  BTRY(b, hipSetDevice(b->device));
  if (int rc = upload(b)) return rc;
  const int M = int(b->mols.size());
  const auto t0 = std::chrono::steady_clock::now();
  auto since = [](std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
  };

  // core-Hamiltonian guess, then the captured iteration until every molecule is done
  launch_step(b, MODE_GUESS, opt->diis, opt->threshold, opt->iterations);
  char key[128];
  snprintf(key, sizeof key, "%d/%d/%d/%a/%d/%p", b->nt, b->chunk, opt->diis, opt->threshold, opt->iterations,
           (void *)b->d_items);
  if (!b->graph || b->graph_key != key) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    ZqCaptureWindow window;
    BTRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_iteration(b, opt->diis, opt->threshold, opt->iterations);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return bfail(b, ZQ_ERR_LAUNCH, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) return bfail(b, ZQ_ERR_LAUNCH, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
    b->graph_key = key;
  }
  hipEvent_t ev[2];
  BTRY(b, hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  BTRY(b, hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  for (int k = 0; k < opt->iterations; ++k) {
    BTRY(b, hipGraphLaunch(b->graph, b->stream));
    BTRY(b, hipMemcpyAsync(b->h_count + (k & 1), b->d_count, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    BTRY(b, hipEventRecord(ev[k & 1], b->stream));
    if (k >= 1) {
      BTRY(b, hipEventSynchronize(ev[(k - 1) & 1]));
      if (b->h_count[(k - 1) & 1] >= M) break;
    }
  }
  BTRY(b, hipStreamSynchronize(b->stream));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  std::vector<BState> hs(M);
  BTRY(b, hipMemcpyAsync(hs.data(), b->d_state, M * sizeof(BState), hipMemcpyDeviceToHost, b->stream));
  BTRY(b, hipStreamSynchronize(b->stream));
  b->stage_ms[0] = since(t0);
  const auto t1 = std::chrono::steady_clock::now();

  // eigenpairs of X^T F_ext X: molecules ordered by n so that each size is one strided batch
  std::vector<int> order(M);
  for (int m = 0; m < M; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->mols[a].n < b->mols[c].n; });
  std::vector<long long> eoff(M), coff(M), woff(M);
  long long etot = 0, wtot = 0;
  for (int m : order) {
    eoff[m] = etot, woff[m] = wtot;
    etot += (long long)b->mols[m].n * b->mols[m].n;
    wtot += b->mols[m].n;
  }
  for (int m = 0; m < M; ++m) coff[m] = eoff[m];
  // the previous run's copies of C and e must have landed before their buffers are reused
  if (b->copy_stream) BTRY(b, hipStreamSynchronize(b->copy_stream));
  if (int rc = eig_buffers(b, size_t(etot), size_t(wtot), size_t(M))) return rc;
  double *d_eig = b->d_eig, *d_w = b->d_w, *d_e = b->d_e, *d_c = b->d_c;
  long long *d_eoff = b->d_eoff;
  int *d_info = b->d_info;
  BTRY(b, hipMemcpyAsync(d_eoff, eoff.data(), M * sizeof(long long), hipMemcpyHostToDevice, b->stream));
  if (b->nt == 512)
    hipLaunchKernelGGL(k_batch_fprime<512>, dim3(M), dim3(512), 0, b->stream, b->ws, b->d_mols, d_eig, d_eoff);
  else
    hipLaunchKernelGGL(k_batch_fprime<256>, dim3(M), dim3(256), 0, b->stream, b->ws, b->d_mols, d_eig, d_eoff);
  BTRY(b, hipGetLastError());
  if (!b->solver) {
    if (rocblas_create_handle(&b->solver) != rocblas_status_success) return bfail(b, ZQ_ERR_LAPACK, "rocblas handle");
  }
  rocblas_set_stream(b->solver, b->stream);
  for (int l = 0; l < zq_batch::kEigLanes; ++l) {
    if (!b->eig_lane[l]) BTRY(b, hipStreamCreateWithFlags(&b->eig_lane[l], hipStreamNonBlocking));
    if (!b->eig_solver[l]) {
      if (rocblas_create_handle(&b->eig_solver[l]) != rocblas_status_success)
        return bfail(b, ZQ_ERR_LAPACK, "rocblas handle");
      rocblas_set_stream(b->eig_solver[l], b->eig_lane[l]);
    }
  }
  for (hipEvent_t &e : b->eig_ev)
    if (!e) BTRY(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // one strided batch per distinct n; the groups go to the least loaded lane (the batch's stream is
  // lane 0), largest first by n^3 x count
  struct Group {
    int s, cnt, n;
    double cost;
  };
  std::vector<Group> groups;
  for (int s = 0; s < M;) {
    const int n = b->mols[order[s]].n;
    int t = s;
    while (t < M && b->mols[order[t]].n == n) ++t;
    groups.push_back(Group{s, t - s, n, double(n) * n * n * (t - s)});
    s = t;
  }
  std::stable_sort(groups.begin(), groups.end(), [](const Group &x, const Group &y) { return x.cost > y.cost; });
  BTRY(b, hipEventRecord(b->eig_ev[0], b->stream));  // F' is ready
  for (int l = 0; l < zq_batch::kEigLanes; ++l) BTRY(b, hipStreamWaitEvent(b->eig_lane[l], b->eig_ev[0], 0));
  double load[zq_batch::kEigLanes + 1] = {};
  for (const Group &g : groups) {
    int lane = 0;
    for (int l = 1; l <= zq_batch::kEigLanes; ++l)
      if (load[l] < load[lane]) lane = l;
    load[lane] += g.cost;
    const int m0 = order[g.s];
    const rocblas_status rs = rocsolver_dsyevd_strided_batched(
        lane == 0 ? b->solver : b->eig_solver[lane - 1], rocblas_evect_original, rocblas_fill_upper, g.n,
        d_eig + eoff[m0], g.n, rocblas_stride(g.n) * g.n, d_w + woff[m0], rocblas_stride(g.n), d_e + woff[m0],
        rocblas_stride(g.n), d_info + g.s, g.cnt);
    if (rs != rocblas_status_success) return bfail(b, ZQ_ERR_LAPACK, "rocsolver_dsyevd_strided_batched");
  }
  // the batch's stream waits for every lane before the eigenvectors are used
  for (int l = 0; l < zq_batch::kEigLanes; ++l) {
    BTRY(b, hipEventRecord(b->eig_ev[1 + l], b->eig_lane[l]));
    BTRY(b, hipStreamWaitEvent(b->stream, b->eig_ev[1 + l], 0));
  }
  if (b->nt == 512)
    hipLaunchKernelGGL(k_batch_coef<512>, dim3(M), dim3(512), 0, b->stream, b->ws, b->d_mols, d_eig, d_eoff, d_c, d_eoff);
  else
    hipLaunchKernelGGL(k_batch_coef<256>, dim3(M), dim3(256), 0, b->stream, b->ws, b->d_mols, d_eig, d_eoff, d_c, d_eoff);
  BTRY(b, hipGetLastError());
  // C, e and the info codes go to pinned host memory on the copy stream, overlapped with the MP2
  BTRY(b, hipEventRecord(b->coef_done, b->stream));
  BTRY(b, hipStreamWaitEvent(b->copy_stream, b->coef_done, 0));
  BTRY(b, hipMemcpyAsync(b->h_c, d_c, etot * sizeof(double), hipMemcpyDeviceToHost, b->copy_stream));
  BTRY(b, hipMemcpyAsync(b->h_w, d_w, wtot * sizeof(double), hipMemcpyDeviceToHost, b->copy_stream));
  BTRY(b, hipMemcpyAsync(b->h_info, d_info, M * sizeof(int), hipMemcpyDeviceToHost, b->copy_stream));
  b->c_off = coff, b->w_off = woff;
  b->res_n.resize(M);
  for (int m = 0; m < M; ++m) b->res_n[m] = b->mols[m].n;
  BTRY(b, hipStreamSynchronize(b->stream));
  b->stage_ms[1] = since(t1);
  const auto t2 = std::chrono::steady_clock::now();

  // MP2 of every molecule that finished its iterations
  std::vector<int> rc_m(M, ZQ_OK);
  for (int m = 0; m < M; ++m) E_corr[m] = 0.0;
  if (mp2_precision != ZQ_PREC_BF16X3) {
    // grouped: one launch per transform stage for all molecules, on the device C and e
    const bool f32 = mp2_precision == ZQ_PREC_FP32 || mp2_precision == ZQ_PREC_FP32C;
    if (f32 && !b->eri32) {
      BTRY(b, hipMalloc(&b->eri32, b->eri_count * sizeof(float)));
      zq_narrow_eri(b->stream, b->eri, b->eri_count, b->eri32);
      BTRY(b, hipGetLastError());
    }
    std::vector<ZqGroupMol> gm;
    std::vector<int> which;
    for (int m = 0; m < M; ++m) {
      if (hs[m].status != 1 && hs[m].status != 2) continue;
      gm.push_back(ZqGroupMol{b->eri + b->eri_off[m], f32 ? b->eri32 + b->eri_off[m] : nullptr, b->mols[m].n,
                              b->mols[m].nocc, d_c + coff[m], d_w + woff[m]});
      which.push_back(m);
    }
    std::vector<double> ec(gm.size(), 0.0);
    std::string err;
    if (int rc = zq_mp2_grouped(b->stream, gm, mp2_precision, &b->mp2_ws, &b->mp2_ws_bytes, ec.data(), err))
      return bfail(b, rc, err);
    for (size_t k = 0; k < which.size(); ++k) E_corr[which[k]] = ec[k];
  } else {
    // BF16X3: several view contexts at once, each with its own stream, on the host copies of C and e
    BTRY(b, hipStreamSynchronize(b->copy_stream));
    const int nworkers = std::min(M, 8);
    std::vector<zq_ctx *> views(nworkers, nullptr);
    for (int w = 0; w < nworkers; ++w) {
      views[w] = zq_create(b->device);
      if (!views[w]) {
        for (zq_ctx *v : views)
          if (v) zq_destroy(v);
        return bfail(b, ZQ_ERR_LAUNCH, "batch MP2: context creation");
      }
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int w = 0; w < nworkers; ++w)
      th.emplace_back([&, w] {
        ZqGate gate;
        (void)hipSetDevice(b->device);
        for (int m; (m = next.fetch_add(1)) < M;) {
          if (hs[m].status != 1 && hs[m].status != 2) continue;
          bind_view(views[w], b, m);
          rc_m[m] = zq_mp2_device(views[w], b->h_c + coff[m], b->h_w + woff[m], b->mols[m].nocc, mp2_precision,
                                  &E_corr[m]);
          views[w]->eri = nullptr;
        }
      });
    for (auto &t : th) t.join();
    for (zq_ctx *v : views) {
      v->eri = nullptr;
      zq_destroy(v);
    }
  }
  b->stage_ms[2] = since(t2);
  BTRY(b, hipStreamSynchronize(b->copy_stream));  // C, e and the info codes are on the host
  std::vector<int> info(M);  // rocSOLVER wrote info by batch position; by molecule here
  for (int s = 0; s < M; ++s) info[order[s]] = b->h_info[s] ? 1 : 0;
  b->stage_ms[3] = since(t0);
  for (int m = 0; m < M; ++m) {
    E_rhf[m] = hs[m].e_new;
    iterations[m] = hs[m].it;
    status[m] = hs[m].status == 1 ? ZQ_OK : hs[m].status == 2 ? ZQ_ERR_NOT_CONVERGED : ZQ_ERR_LAPACK;
    if (info[m] && status[m] == ZQ_OK) status[m] = ZQ_ERR_LAPACK;
    if (rc_m[m] != ZQ_OK && status[m] == ZQ_OK) status[m] = rc_m[m];
  }
  return ZQ_OK;
}

int zq_batch_timings(const zq_batch *b, double ms[4]) {
  if (!b || !ms) return ZQ_ERR_INVALID;
  std::copy(b->stage_ms, b->stage_ms + 4, ms);
  return ZQ_OK;
}

int zq_batch_orbitals(const zq_batch *b, int index, double *e, double *C) {
  if (!b || index < 0 || size_t(index) >= b->res_n.size() || !e || !C || !b->h_c || !b->h_w) return ZQ_ERR_INVALID;
  const long long n = b->res_n[index];
  std::copy(b->h_w + b->w_off[index], b->h_w + b->w_off[index] + n, e);
  std::copy(b->h_c + b->c_off[index], b->h_c + b->c_off[index] + n * n, C);
  return ZQ_OK;
}

}  // extern "C"
