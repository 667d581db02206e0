// Device-resident SCF: the loop of src/hartree_fock.zig:738-875 with every n x n step on the GPU.
//
//   per iteration (all on the context's stream, FP64):
//     F      = H + G(P)                           J/K kernels (getFock, :686-735)
//     E      = 1/2 sum P (H + F) + V_NN           per-block partials from the J/K reduce, summed in a
//                                                 fixed order (getEnergy, :643-655)
//     err    = F P S - (F P S)^T                   MFMA GEMMs (getError, :658-683)
//     B row  = <err_new, err_i>                   deterministic reductions; the <= 9 x 9 DIIS system
//                                                 is solved by one wave (diis, :176-227) and the
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
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <vector>

#include "../../include/zinq_scf.h"
#include "zq_internal.h"
#include "zq_linalg.h"
#include "zq_tile.h"

namespace {

constexpr int kRedBlocks = 64;

// partial[b] = sum over a fixed strided subset of elements of x[i]*y[i] (or of terms), deterministic.
__global__ __launch_bounds__(256) void k_energy_part(const double *__restrict__ P, int npad, const double *__restrict__ H,
                                                     const double *__restrict__ F, int n, double *__restrict__ part,
                                                     const int *__restrict__ skip) {
  __shared__ double red[256];
  if (skip && *skip) return;
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

// err = scale (FPS - (FPS)^T); scale is 1 outside the singular-system test hook (exact: x * 1 = x)
__global__ void k_antisym(const double *__restrict__ FPS, int n, double scale, double *__restrict__ err) {
  for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < size_t(n) * n; e += size_t(gridDim.x) * blockDim.x) {
    const int i = int(e / n), j = int(e % n);
    err[e] = scale * (FPS[e] - FPS[size_t(j) * n + i]);
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

// P[i][j] (padded npad stride) = Pn[i][j]; part = sum (Pn - P_old)^2; dP (if given) = Pn - P_old.
__global__ __launch_bounds__(256) void k_store_density(const double *__restrict__ Pn, int n, double *__restrict__ P, int npad,
                                                       double *__restrict__ part, const int *__restrict__ skip,
                                                       double *__restrict__ dP) {
  __shared__ double red[256];
  if (skip && *skip) return;
  double s = 0.0;
  for (size_t e = blockIdx.x * 256 + threadIdx.x; e < size_t(n) * n; e += size_t(gridDim.x) * 256) {
    const int i = int(e / n), j = int(e % n);
    const double v = Pn[e];
    const double d = v - P[size_t(i) * npad + j];
    s += d * d;
    P[size_t(i) * npad + j] = v;
    if (dP) dP[size_t(i) * npad + j] = d;
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
  double *dP = nullptr;     // P_k - P_(k-1), npad x npad zero padded (mixed Fock path only)
  double *epart = nullptr;  // energy partials of the J/K reduce, one per output block
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
              {"scf_Eh", &d.Eh, nn * maxh}, {"scf_part", &d.part, size_t(kRedBlocks) * (2 + maxh)},
              {"scf_scal", &d.scal, size_t(2 + maxh)}, {"scf_coef", &d.coef, size_t(2 + maxh)}, {"scf_eig", &d.eig, size_t(n)},
              {"scf_offd", &d.offd, size_t(n)}};
  for (auto &b : bufs) {
    *b.p = (double *)zq_scratch(ctx, b.name, b.cnt * sizeof(double));
    if (!*b.p) return zq_fail(ctx, ZQ_ERR_OOM, std::string("device SCF buffer ") + b.name);
  }
  d.slots = (int *)zq_scratch(ctx, "scf_slots", 64 * sizeof(int));
  d.info = (int *)zq_scratch(ctx, "scf_info", 4 * sizeof(int));
  if (!d.slots || !d.info) return zq_fail(ctx, ZQ_ERR_OOM, "device SCF buffers");
  // the final eigensolve's buffers (zq_eigen.hip) too: a buffer first allocated after the iteration was
  // captured would empty the capture cache at the next run
  if (!zq_scratch(ctx, "eig_Q", nn * sizeof(double)) || !zq_scratch(ctx, "eig_Z", nn * sizeof(double)))
    return zq_fail(ctx, ZQ_ERR_OOM, "eigen-solver buffers");
  return ZQ_OK;
}

// Fp = X^T A X for the symmetric n x n A (X^T stored as XT, so both products are A B^T forms).
int to_orthonormal(Dev &d, const double *A) {
  const int n = d.n;
  if (int rc = zq_gemm_nt(d.ctx, d.XT, n, A, n, d.T, n, n, n, n, 1.0, "scf_gemm")) return rc;  // T = X^T A
  return zq_gemm_nt(d.ctx, d.T, n, d.XT, n, d.Fp, n, n, n, n, 1.0, "scf_gemm");                 // Fp = T X
}

// Eigen-decomposition of Fp (symmetric) by zq_syev_rows, C = X V, eigenvalues ascending in d.eig.
// Fp is overwritten by the eigenvectors.
int diag_orthonormal(Dev &d, double *Cout) {
  zq_ctx *ctx = d.ctx;
  const int n = d.n;
  if (int rc = zq_syev_rows(ctx, d.Fp, n, d.eig, d.offd, d.info)) return rc;
  // eigenvector k is row k: C[i][k] = sum_p X[i][p] V[k][p]
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
    hipLaunchKernelGGL(k_store_density, dim3(kRedBlocks), dim3(256), 0, ctx->stream, d.Pn, n, ctx->d_P, np, d.part,
                       nullptr, nullptr);
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

// ---------------------------------------------------------------------------------------------------
// Asynchronous driver: the whole iteration, DIIS bookkeeping included, is decided on the device.  Every
// kernel of an iteration returns at once when state.done != 0, so the host can enqueue iteration k+1
// before it has seen whether iteration k converged; it reads the state one iteration late.

constexpr int kMaxHist = 16;

struct ScfState {
  int done;       // 0 running, 1 converged, 2 iteration limit reached, 3 SP2 did not settle in its budget
  int it;         // iterations completed
  int sp2_steps;  // SP2 steps of the last density
  int sp2_max;    // most SP2 steps of any density so far
  int hist_n, cur, combine;
  int resets;          // singular DIIS systems met (the history reset to the current entry)
  int hist[kMaxHist];  // DIIS slots, oldest first
  int sp2[6];          // SP2 {count, steps} pairs, the settled buffer's parity, the abort flag
  double e_old, e_new, rms, dE;
  double coef[kMaxHist + 1];
  double B[kMaxHist * kMaxHist];  // B[slot][slot] = <err_slot, err_slot'>
};

// Slot of the next iteration's error/Fock pair: the oldest when the history is full, else the lowest
// free.  Taken at the end of the previous iteration (k_converge) or by k_state_init for the first.
__device__ void diis_next_slot(ScfState *st, int diis) {
  int slot;
  if (st->hist_n >= diis) {
    slot = st->hist[0];
    for (int k = 1; k < st->hist_n; ++k) st->hist[k - 1] = st->hist[k];
    --st->hist_n;
  } else {
    slot = 0;
    for (bool used = true; used; ) {
      used = false;
      for (int k = 0; k < st->hist_n; ++k) used |= st->hist[k] == slot;
      if (used) ++slot;
    }
  }
  st->hist[st->hist_n++] = slot;
  st->cur = slot;
}

__global__ void k_state_init(ScfState *st, double VN, int diis) {
  if (threadIdx.x != 0) return;
  st->done = 0, st->it = 0, st->sp2_steps = 0, st->sp2_max = 0;
  st->hist_n = 0, st->cur = 0, st->combine = 0, st->resets = 0;
  st->e_old = VN, st->e_new = VN, st->rms = 0.0, st->dE = 0.0;
  if (diis > 0) diis_next_slot(st, diis);
}

// err = scale (FPS - (FPS)^T) into Eh[cur], F into Fh[cur].
__global__ void k_antisym_store(const double *__restrict__ FPS, const double *__restrict__ F, int n, double scale,
                                double *__restrict__ Eh, double *__restrict__ Fh, const ScfState *__restrict__ st) {
  if (st->done) return;
  const size_t nn = size_t(n) * n, off = size_t(st->cur) * nn;
  for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < nn; e += size_t(gridDim.x) * blockDim.x) {
    const int i = int(e / n), j = int(e % n);
    Eh[off + e] = scale * (FPS[e] - FPS[size_t(j) * n + i]);
    Fh[off + e] = F[e];
  }
}

// part[slot*kRedBlocks + b]: dot(err_cur, err_slot) over a strided subset (blockIdx.y = slot).
__global__ __launch_bounds__(256) void k_dots_state(const double *__restrict__ hist, size_t nn, double *__restrict__ part,
                                                    const ScfState *__restrict__ st) {
  __shared__ double red[256];
  if (st->done) return;
  const double *enew = hist + size_t(st->cur) * nn, *ei = hist + size_t(blockIdx.y) * nn;
  double s = 0.0;
  if (int(blockIdx.y) < kMaxHist) {
    bool live = false;
    for (int k = 0; k < st->hist_n; ++k) live |= st->hist[k] == int(blockIdx.y);
    if (live)
      for (size_t e = blockIdx.x * 256 + threadIdx.x; e < nn; e += size_t(gridDim.x) * 256) s += enew[e] * ei[e];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.y * kRedBlocks + blockIdx.x] = red[0];
}

// LU with partial pivoting and the triangular solves of the host zq::getrf/getrs, on one wave: row i
// of A (n <= kLuMax) in LDS at A[i * kLuLd], b[i] in lane i's register.  Every element sees the same
// operations in the same order as in the serial steps: the same pivot (the first row with the largest
// |A[i][k]|), the same multipliers and updates, the forward sums in ascending j, the backward sums in
// ascending j.  The rows of b are swapped as the pivots are found, which permutes b exactly as applying
// the swaps afterwards does.  The loops run to n, not to kLuMax: the code stays small, which matters for
// a kernel that runs once per iteration with its instructions not in the cache.  Returns false, on every
// lane, for a zero pivot; otherwise lane i < n holds x[i] in b.
constexpr int kLuMax = kMaxHist + 1;
constexpr int kLuLd = kLuMax + 1;  // row stride of A in LDS

// lane src's v on every lane; src is wave-uniform, so this is two v_readlane, not an LDS permute
__device__ inline double lane_bcast(double v, int src) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), src),
                          __builtin_amdgcn_readlane(__double2loint(v), src));
}

__device__ bool lu_solve_lds(int n, double *A, double &b) {
  const int lane = threadIdx.x & 63;
  for (int k = 0; k < n; ++k) {
    // pivot: every lane scans column k in row order, so all find the same row
    int p = k;
    double best = fabs(A[k * kLuLd + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = fabs(A[i * kLuLd + k]);
      if (v > best) best = v, p = i;
    }
    if (best == 0.0) return false;
    p = __builtin_amdgcn_readfirstlane(p);
    if (p != k) {
      if (lane < n) {  // lane j swaps column j of rows k and p
        const double rk = A[k * kLuLd + lane], rp = A[p * kLuLd + lane];
        A[k * kLuLd + lane] = rp;
        A[p * kLuLd + lane] = rk;
      }
      const double bk = lane_bcast(b, k), bp = lane_bcast(b, p);
      b = lane == k ? bp : (lane == p ? bk : b);
      __syncthreads();
    }
    if (lane > k && lane < n) {
      const double f = A[lane * kLuLd + k] / A[k * kLuLd + k];
      A[lane * kLuLd + k] = f;
      for (int j = k + 1; j < n; ++j) A[lane * kLuLd + j] -= f * A[k * kLuLd + j];
    }
    __syncthreads();
  }
  for (int j = 0; j < n; ++j) {
    const double bj = lane_bcast(b, j);
    if (lane > j && lane < n) b -= A[lane * kLuLd + j] * bj;
  }
  for (int i = n - 1; i >= 0; --i) {
    double acc = b;
    for (int j = i + 1; j < n; ++j) {
      const double bj = lane_bcast(b, j);
      if (lane < n) acc -= A[lane * kLuLd + j] * bj;
    }
    if (lane == i) b = acc / A[i * kLuLd + i];
  }
  return true;
}

// E = sum of the energy partials + V_NN; DIIS: the new row of B, the bordered system, its LU solve
// (src/hartree_fock.zig:176-227), or a reset to the current entry when it is singular (:821-845).
// One wave: lane b holds partial b of every sum (kRedBlocks == 64), each sum a fixed-order butterfly;
// the wave solves the <= 17 x 17 system, one row per lane.
// epart: the nep energy partials the J/K reduce wrote (one per output block).
__global__ __launch_bounds__(64) void k_diis_solve(const double *__restrict__ part, const double *__restrict__ epart,
                                                   int nep, int diis, double VN, ScfState *__restrict__ st) {
  static_assert(kRedBlocks == 64, "one partial per lane");
  __shared__ double Bs[kMaxHist * kMaxHist];
  if (st->done) return;
  const int lane = threadIdx.x;
  const int m = st->hist_n, slot = st->cur;
  // every load is issued before any of them is used: the history, all the dot partials, the energy
  int hist[kMaxHist];  // the history in registers: every lane needs it
  double dp[kMaxHist];
#pragma unroll
  for (int h = 0; h < kMaxHist; ++h) hist[h] = st->hist[h];
  const int hist_lane = st->hist[min(lane, kMaxHist - 1)];
#pragma unroll
  for (int k = 0; k < kMaxHist; ++k) dp[k] = part[(1 + k) * kRedBlocks + lane];
  double e = 0.0;
  for (int b = lane; b < nep; b += 64) e += epart[b];
  // the live slots as a mask, the history and the dot products in LDS (indexed at run time below)
  __shared__ int hs[kMaxHist];
  __shared__ double dots[kMaxHist];
  unsigned live = 0;
#pragma unroll
  for (int h = 0; h < kMaxHist; ++h) live |= h < m ? 1u << hist[h] : 0u;
  if (lane < kMaxHist) hs[lane] = lane < m ? hist_lane : -1;
  e = zqt::wave_sum(e);
#pragma unroll
  for (int k = 0; k < kMaxHist; ++k) {
    double d = 0.0;
    if (diis > 0 && (live >> k & 1)) d = zqt::wave_sum(dp[k]);  // uniform branch
    if (lane == 0) dots[k] = d;
  }
  __syncthreads();
  // the m x m block of B over the history, one entry per lane, with the new row and column from dots
  for (int idx = lane; diis > 0 && m >= 2 && idx < m * m; idx += 64) {
    const int i = idx / m, j = idx % m;
    const int si = hs[i], sj = hs[j];
    Bs[i * kMaxHist + j] = (si == slot || sj == slot) ? dots[si == slot ? sj : si] : st->B[si * kMaxHist + sj];
  }
  __syncthreads();
  if (lane == 0) {
    st->e_new = e + VN;
    st->combine = 0;
    if (diis > 0)
      for (int k = 0; k < kMaxHist; ++k)
        if (live >> k & 1) st->B[slot * kMaxHist + k] = st->B[k * kMaxHist + slot] = dots[k];
  }
  if (diis <= 0 || m < 2) return;
  // the bordered system [B -1; -1 0] c = [0; -1], row i written by lane i
  __shared__ double Alu[kLuMax * kLuLd];
  if (lane <= m)
    for (int j = 0; j <= m; ++j)
      Alu[lane * kLuLd + j] = lane < m ? (j < m ? Bs[lane * kMaxHist + j] : -1.0) : (j < m ? -1.0 : 0.0);
  double c = lane == m ? -1.0 : 0.0;
  __syncthreads();
  if (lu_solve_lds(m + 1, Alu, c)) {
    if (lane < m) st->coef[lane] = c;
    if (lane == 0) st->combine = 1;
  } else if (lane == 0) {
    st->hist[0] = slot;
    st->hist_n = 1;
    ++st->resets;
  }
}

// W = sum_i c_i Fh[hist_i] symmetrised when DIIS combined, else W = F.
__global__ void k_combine_state(const double *__restrict__ Fh, const double *__restrict__ F, int n, double *__restrict__ W,
                                const ScfState *__restrict__ st) {
  if (st->done) return;
  const size_t nn = size_t(n) * n;
  const int m = st->combine ? st->hist_n : 0;
  for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < nn; e += size_t(gridDim.x) * blockDim.x) {
    const int i = int(e / n), j = int(e % n);
    if (m == 0) {
      W[e] = F[e];
      continue;
    }
    if (j < i) continue;
    double a = 0.0, b = 0.0;
    for (int k = 0; k < m; ++k) {
      const double *Fk = Fh + size_t(st->hist[k]) * nn;
      a += st->coef[k] * Fk[size_t(i) * n + j];
      b += st->coef[k] * Fk[size_t(j) * n + i];
    }
    const double avg = (a + b) / 2;
    W[size_t(i) * n + j] = avg;
    W[size_t(j) * n + i] = avg;
  }
}

// The SP2 density of this iteration: steps taken, and done = 3 (the caller falls back to the
// host-checked driver) when it did not settle in its budget or a bounded wait gave up.
__global__ void k_sp2_check(int parity, ScfState *__restrict__ st) {
  if (threadIdx.x != 0 || st->done) return;
  const int *fin = st->sp2 + 2 * parity;
  st->sp2_steps = fin[1];
  st->sp2_max = max(st->sp2_max, fin[1]);
  if (fin[0] < 2 || st->sp2[5] != 0) st->done = 3;
}

// The end of an iteration: the SP2 check of k_sp2_check (parity of its step budget), then rms(dP) and
// the convergence test of src/hartree_fock.zig:853-870 (one wave, butterfly sum).  The state's done
// word also goes to *host_done (host memory the device writes), which the host reads one iteration late.
__global__ __launch_bounds__(64) void k_converge(const double *__restrict__ part, int n, double thr, int maxit,
                                                 ScfState *__restrict__ st, int parity, int *__restrict__ host_done,
                                                 int diis) {
  if (st->done) {
    // set before this iteration (the guess density's SP2 check): the host stops replaying as well
    if (threadIdx.x == 0) *host_done = st->done;
    return;
  }
  const double s = zqt::wave_sum(part[threadIdx.x]);
  if (threadIdx.x != 0) return;
  const int *fin = st->sp2 + 2 * parity;
  st->sp2_steps = fin[1];
  st->sp2_max = max(st->sp2_max, fin[1]);
  if (fin[0] < 2 || st->sp2[5] != 0) {  // SP2 did not settle: the P of this iteration is not used
    st->done = 3;
    *host_done = 3;
    return;
  }
  const double rms = sqrt(s / (double(n) * n));
  const double dE = fabs(st->e_new - st->e_old);
  st->e_old = st->e_new;
  st->rms = rms, st->dE = dE;
  const bool conv = st->it > 0 && dE < thr && rms < thr;
  ++st->it;
  if (conv)
    st->done = 1;
  else if (st->it >= maxit)
    st->done = 2;
  else if (diis > 0)
    diis_next_slot(st, diis);
  *host_done = st->done;
}

// P from the symmetric Fock matrix A by SP2 with a fixed step budget, without host synchronisation.
// check: k_sp2_check follows SP2 (else the iteration's k_converge makes the check).
int density_async(Dev &d, const double *A, int nocc, double factor, int nsteps, ScfState *st, double *part_rms,
                  double *dP, bool check) {
  zq_ctx *ctx = d.ctx;
  const int n = d.n, np = ctx->npad;
  const int *skip = &st->done;
  if (int rc = zq_gemm_nt(ctx, d.XT, n, A, n, d.T, n, n, n, n, 1.0, "scf_gemm", skip)) return rc;
  if (int rc = zq_gemm_nt(ctx, d.T, n, d.XT, n, d.Fp, n, n, n, n, 1.0, "scf_gemm", skip)) return rc;
  const double *D = nullptr;
  long long stride = 0;
  int ld = 0;
  // the first density of the SCF starts the fold bookkeeping; later ones may fold
  if (int rc = zq_sp2_enqueue(ctx, d.Fp, n, nocc, 1e-9, nsteps, skip, st->sp2, &D, &stride, &ld, check ? 1 : 2)) return rc;
  if (check) hipLaunchKernelGGL(k_sp2_check, dim3(1), dim3(64), 0, ctx->stream, nsteps & 1, st);
  if (int rc = zq_gemm_nt(ctx, d.X, n, D, ld, d.T, n, n, n, n, 1.0, "scf_gemm", skip, st->sp2 + 4, stride)) return rc;
  if (int rc = zq_gemm_nt(ctx, d.T, n, d.X, n, d.Pn, n, n, n, n, factor, "scf_gemm", skip)) return rc;
  hipLaunchKernelGGL(k_store_density, dim3(kRedBlocks), dim3(256), 0, ctx->stream, d.Pn, n, ctx->d_P, np, part_rms, skip,
                     dP);
  ZQ_LAUNCH_CHECK(ctx, "density_async");
  return ZQ_OK;
}

// Destroys the context's captured iterations.
void drop_graphs(zq_ctx *ctx) {
  for (auto &kv : ctx->scf_graphs)
    if (kv.second) (void)hipGraphExecDestroy((hipGraphExec_t)kv.second);
  ctx->scf_graphs.clear();
}

struct Run {
  int n, nocc, np;
  double x_exch, factor, VN;
  const zq_scf_options *opt;
  bool mixed;  // incremental Fock builds from the FP32 ERI copy
};

// Uploads S, X = L^-T, X^T, H; zeroes P.
int setup(Dev &d, const double *H, const double *S) {
  zq_ctx *ctx = d.ctx;
  const int n = d.n, np = ctx->npad;
  const size_t nn = d.nn;
  std::vector<double> hX(nn), hXT(nn);
  if (zq::chol_inv_t(n, S, hX.data())) return zq_fail(ctx, ZQ_ERR_LAPACK, "overlap matrix not positive definite");
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) hXT[size_t(j) * n + i] = hX[size_t(i) * n + j];
  ZQ_TRY(ctx, hipMemcpyAsync(d.S, S, nn * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(d.X, hX.data(), nn * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(d.XT, hXT.data(), nn * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(ctx->d_H, H, nn * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemsetAsync(ctx->d_P, 0, size_t(np) * np * 8, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));  // the host vectors go out of scope
  return ZQ_OK;
}

// The kernels of one SCF iteration; every one of them returns at once while state.done != 0.
int enqueue_iteration(Dev &d, const Run &r, int budget, ScfState *st, double *part_rms, int *host_done) {
  zq_ctx *ctx = d.ctx;
  const int n = d.n, maxh = r.opt->diis > 0 ? r.opt->diis : 1;
  const size_t nn = d.nn;
  const int gblocks = int(std::min<size_t>(1024, (nn + 255) / 256));
  const int *skip = &st->done;
  // the reduce also leaves the energy partials (getEnergy) in d.epart; the DIIS slot of this iteration
  // was taken by the previous one's k_converge (or k_state_init)
  if (r.mixed) {
    // F_k = F_(k-1) + G(P_k - P_(k-1)), in place on d_F, from the FP32 ERI copy
    if (int rc = zq_fock_jk_device(ctx, d.dP, ctx->d_F, r.x_exch, ctx->d_F, skip, true, d.epart)) return rc;
  } else if (int rc = zq_fock_jk_device(ctx, ctx->d_P, ctx->d_H, r.x_exch, ctx->d_F, skip, false, d.epart)) {
    return rc;
  }
  if (r.opt->diis > 0) {
    // F P (P symmetric) and (F P) S (S symmetric) as A B^T products
    if (int rc = zq_gemm_nt(ctx, ctx->d_F, n, ctx->d_P, r.np, d.FP, n, n, n, n, 1.0, "scf_gemm", skip)) return rc;
    if (int rc = zq_gemm_nt(ctx, d.FP, n, d.S, n, d.FPS, n, n, n, n, 1.0, "scf_gemm", skip)) return rc;
    hipLaunchKernelGGL(k_antisym_store, dim3(gblocks), dim3(256), 0, ctx->stream, d.FPS, ctx->d_F, n, ctx->diis_err_scale,
                       d.Eh, d.Fh, st);
    hipLaunchKernelGGL(k_dots_state, dim3(kRedBlocks, maxh), dim3(256), 0, ctx->stream, d.Eh, nn, d.part + kRedBlocks, st);
  }
  hipLaunchKernelGGL(k_diis_solve, dim3(1), dim3(64), 0, ctx->stream, d.part, d.epart, ctx->Nb, r.opt->diis, r.VN, st);
  hipLaunchKernelGGL(k_combine_state, dim3(gblocks), dim3(256), 0, ctx->stream, d.Fh, ctx->d_F, n, d.W, st);
  if (int rc = density_async(d, d.W, r.nocc, r.factor, budget, st, part_rms, r.mixed ? d.dP : nullptr, false)) return rc;
  hipLaunchKernelGGL(k_converge, dim3(1), dim3(64), 0, ctx->stream, part_rms, n, r.opt->threshold, r.opt->iterations, st,
                     budget & 1, host_done, r.opt->diis);
  ZQ_LAUNCH_CHECK(ctx, "scf iteration");
  return ZQ_OK;
}

// The asynchronous SCF.  Returns ZQ_OK with *status = state.done (1 converged, 2 iteration limit,
// 3 SP2 budget exceeded: the caller reruns the synchronised driver), the final energy and iteration
// count, and W = the last diagonalised Fock matrix.
int scf_async(Dev &d, const Run &r, int *status, double *energy, int *iterations) {
  zq_ctx *ctx = d.ctx;
  const int n = d.n, nocc = r.nocc;
  const size_t nn = d.nn;
  std::unique_ptr<ZqGate> gate(new ZqGate);  // set-up allocates; released before the replay loop
  auto *st = (ScfState *)zq_scratch(ctx, "scf_state", sizeof(ScfState));
  auto *hdone = (int *)zq_pinned(ctx, "scf_done", sizeof(int));
  if (!st || !hdone || !zq_pinned(ctx, "scf_state_full", sizeof(ScfState))) return zq_fail(ctx, ZQ_ERR_OOM, "SCF state");
  int *ddone = nullptr;  // the same host word as the kernels address it
  ZQ_TRY(ctx, hipHostGetDevicePointer((void **)&ddone, hdone, 0));
  double *part_rms = d.part + (1 + kMaxHist) * kRedBlocks;
  d.epart = (double *)zq_scratch(ctx, "scf_epart", size_t(ctx->Nb) * sizeof(double));
  if (!d.epart) return zq_fail(ctx, ZQ_ERR_OOM, "SCF energy partials");
  hipEvent_t *ev = ctx->scf_events;
  if (!ev[0]) ZQ_TRY(ctx, hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  if (!ev[1]) ZQ_TRY(ctx, hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));

  if (r.mixed) {
    if (int rc = zq_eri32(ctx)) return rc;
    d.dP = (double *)zq_scratch(ctx, "scf_dP", size_t(r.np) * r.np * sizeof(double));
    if (!d.dP) return zq_fail(ctx, ZQ_ERR_OOM, "SCF dP");
    ZQ_TRY(ctx, hipMemsetAsync(d.dP, 0, size_t(r.np) * r.np * sizeof(double), ctx->stream));
  }
  hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, ctx->stream, st, r.VN, r.opt->diis);
  // core-Hamiltonian guess, src/hartree_fock.zig:425-431; W = H until the first iteration replaces it
  ZQ_TRY(ctx, hipMemcpyAsync(d.W, ctx->d_H, nn * 8, hipMemcpyDeviceToDevice, ctx->stream));
  if (int rc = density_async(d, ctx->d_H, nocc, r.factor, ctx->sp2_guess_budget, st, part_rms,
                             r.mixed ? d.dP : nullptr, true))
    return rc;
  // the incremental builds start from F = H (dP = P_0 after the guess)
  if (r.mixed) ZQ_TRY(ctx, hipMemcpyAsync(ctx->d_F, ctx->d_H, nn * 8, hipMemcpyDeviceToDevice, ctx->stream));
  const int budget = ctx->sp2_budget;

  gate.reset();
  // one iteration as a hipGraph, captured once per context and key, replayed with one launch
  const bool use_graph = !ctx->profile;
  char key[256];
  snprintf(key, sizeof key, "%d/%d/%d/%d/%d/%a/%a/%a/%a/%a/%c%llu/%p/%p/%d", n, nocc, budget, r.opt->diis,
           r.opt->iterations, r.x_exch, r.factor, r.VN, r.opt->threshold, ctx->diis_err_scale, ctx->set_id ? 's' : 'g',
           ctx->set_id ? ctx->set_id : (unsigned long long)ctx->eri_gen, (void *)st, (void *)ddone, int(r.mixed));
  if (ctx->scf_graphs_gen != ctx->scratch_gen || ctx->scf_graphs.size() >= 64) {
    drop_graphs(ctx);
    ctx->scf_graphs_gen = ctx->scratch_gen;
  }
  auto it = ctx->scf_graphs.find(key);
  hipGraphExec_t graph_exec = it == ctx->scf_graphs.end() ? nullptr : (hipGraphExec_t)it->second;
  if (use_graph && !graph_exec) {
    ZqCaptureWindow window;  // no allocation or synchronous call in any thread while capturing
    ZQ_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_iteration(d, r, budget, st, part_rms, ddone);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return zq_check(ctx, e, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) return zq_check(ctx, e2, "hipGraphInstantiate");
    ctx->scf_graphs[key] = exec;
    ++ctx->scf_captures;
    graph_exec = exec;
  }
  // the stream is idle here (the previous run ended with a synchronisation), so no kernel writes it now
  *(volatile int *)hdone = 0;
  ctx->scf_replays = 0;
  int k = 0;
  for (; k < r.opt->iterations; ++k) {
    ++ctx->scf_replays;
    if (use_graph) {
      ZQ_TRY(ctx, hipGraphLaunch(graph_exec, ctx->stream));
    } else if (int rc = enqueue_iteration(d, r, budget, st, part_rms, ddone)) {
      return rc;
    }
    ZQ_TRY(ctx, hipEventRecord(ev[k & 1], ctx->stream));
    if (k >= 1) {
      // done after iteration k-1 (or a later one: done never goes back to 0); iteration k is already
      // queued behind it
      ZQ_TRY(ctx, hipEventSynchronize(ev[(k - 1) & 1]));
      if (*(volatile int *)hdone != 0) break;
    }
  }
  auto *hs_p = (ScfState *)zq_pinned(ctx, "scf_state_full", sizeof(ScfState));
  if (!hs_p) return zq_fail(ctx, ZQ_ERR_OOM, "SCF state");
  ZQ_TRY(ctx, hipMemcpyAsync(hs_p, st, sizeof(ScfState), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const ScfState hs = *hs_p;
  *status = hs.done == 0 ? 2 : hs.done;
  *energy = hs.e_new;
  *iterations = hs.it;
  ctx->scf_diis_resets = hs.resets;
  return ZQ_OK;
}

// The host-synchronised driver: SP2 with host convergence checks and a dsyevd fallback per density,
// DIIS solved on the host.  Used when the asynchronous driver's SP2 budget was too small.
int scf_synced(Dev &d, const Run &r, int *status, double *energy, int *iterations) {
  zq_ctx *ctx = d.ctx;
  const zq_scf_options *opt = r.opt;
  const int n = d.n, nocc = r.nocc, np = r.np, maxh = opt->diis > 0 ? opt->diis : 1;
  const size_t nn = d.nn;
  const int gblocks = int(std::min<size_t>(1024, (nn + 255) / 256));
  const double VN = r.VN;
  ZQ_TRY(ctx, hipMemcpyAsync(d.W, ctx->d_H, nn * 8, hipMemcpyDeviceToDevice, ctx->stream));
  if (int rc = new_density(d, ctx->d_H, nocc, r.factor)) return rc;

  std::deque<int> hist;  // ring-buffer slots of the DIIS history, oldest first
  std::vector<double> Bfull(64 * 64, 0.0);  // B[slot][slot]
  double e_old = VN, e_new = VN;
  int it;
  bool converged = false;
  ctx->scf_diis_resets = 0;
  for (it = 0; it < opt->iterations; ++it) {
    if (int rc = zq_fock_jk_device(ctx, ctx->d_P, ctx->d_H, r.x_exch, ctx->d_F)) return rc;
    hipLaunchKernelGGL(k_energy_part, dim3(kRedBlocks), dim3(256), 0, ctx->stream, ctx->d_P, np, ctx->d_H, ctx->d_F, n,
                       d.part, nullptr);
    const double *Fdiag = ctx->d_F;
    double hbuf[1 + 64];
    if (opt->diis > 0) {
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
      hipLaunchKernelGGL(k_antisym, dim3(gblocks), dim3(256), 0, ctx->stream, d.FPS, n, ctx->diis_err_scale,
                         d.Eh + size_t(slot) * nn);
      ZQ_TRY(ctx, hipMemcpyAsync(d.Fh + size_t(slot) * nn, ctx->d_F, nn * 8, hipMemcpyDeviceToDevice, ctx->stream));
      hist.push_back(slot);
      const int m = int(hist.size());
      hipLaunchKernelGGL(k_dots, dim3(kRedBlocks, maxh), dim3(256), 0, ctx->stream, d.Eh, d.Eh + size_t(slot) * nn, nn,
                         d.part + kRedBlocks);
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(128), 0, ctx->stream, d.part, 1 + maxh, d.scal);
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
          ++ctx->scf_diis_resets;
        }
      }
    } else {
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, ctx->stream, d.part, 1, d.scal);
      if (int rc = sync_read(ctx, d.scal, hbuf, 1)) return rc;
      e_new = hbuf[0] + VN;
    }
    if (Fdiag != d.W) ZQ_TRY(ctx, hipMemcpyAsync(d.W, Fdiag, nn * 8, hipMemcpyDeviceToDevice, ctx->stream));
    if (int rc = new_density(d, d.W, nocc, r.factor)) return rc;
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
  *status = converged ? 1 : 2;
  *energy = e_new;
  *iterations = converged ? it : opt->iterations;
  return ZQ_OK;
}

}  // namespace

void zq_release_solver(zq_ctx *ctx) {
  if (!ctx) return;
  if (ctx->solver) {
    rocblas_destroy_handle((rocblas_handle)ctx->solver);
    ctx->solver = nullptr;
  }
  drop_graphs(ctx);
  for (hipEvent_t &e : ctx->scf_events)
    if (e) (void)hipEventDestroy(e), e = nullptr;
  for (hipEvent_t &e : ctx->eig_events)
    if (e) (void)hipEventDestroy(e), e = nullptr;
}

extern "C" int zq_scf_run_device(zq_ctx *ctx, size_t n_, size_t nocc_, const double *H, const double *S, double VN,
                                 const zq_scf_options *opt, double *C, double *e, double *P, double *F,
                                 zq_scf_result *res) {
  if (!ctx || !H || !S || !opt || !C || !e || !P || !F || !res || n_ == 0 || nocc_ > n_)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_scf_run_device: bad arguments");
  if (!ctx->eri || size_t(ctx->n) != n_) return zq_fail(ctx, ZQ_ERR_STATE, "zq_scf_run_device: ERIs of this size not loaded");
  if (opt->diis > 64) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_scf_run_device: DIIS history above 64");
  // Allocation, synchronous copies and rocSOLVER run under the shared gate; the graph-replay loop of
  // scf_async does not, so another thread may capture while this one iterates.
  std::unique_ptr<ZqGate> gate(new ZqGate);
  if (opt->fock_precision != ZQ_PREC_FP64 && opt->fock_precision != ZQ_PREC_FP32)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_scf_run_device: fock_precision must be ZQ_PREC_FP64 or ZQ_PREC_FP32");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  ctx->res_C = ctx->res_e = nullptr;  // until this run succeeds
  const int n = int(n_), nocc = int(nocc_), np = ctx->npad;
  const bool generalized = opt->generalized != 0;
  Run r{n, nocc, np, generalized ? 1.0 : 0.5, generalized ? 1.0 : 2.0, VN, opt, opt->fock_precision == ZQ_PREC_FP32};
  const int maxh = opt->diis > 0 ? opt->diis : 1;
  Dev d;
  if (int rc = alloc(ctx, d, n, std::max(maxh, kMaxHist))) return rc;
  const size_t nn = d.nn;

  int status = 0, iterations = 0;
  double energy = VN;
  if (int rc = setup(d, H, S)) return rc;
  ctx->scf_replays = 0;
  ctx->scf_diis_resets = 0;
  if (opt->diis <= kMaxHist) {
    gate.reset();
    const int rc = scf_async(d, r, &status, &energy, &iterations);
    gate.reset(new ZqGate);
    if (rc) return rc;
  }
  if (status == 3) ctx->sp2_budget = std::min(kSp2MaxSteps, ctx->sp2_budget + 16);
  ctx->scf_driver = (status == 0 || status == 3) ? 2 : 1;
  ctx->scf_sp2[0] = ctx->scf_sp2[1] = ctx->scf_sp2[2] = 0;
  if (ctx->scf_driver == 1)
    if (int rc = zq_sp2_tracked_stats(ctx, ctx->scf_sp2)) return rc;
  if (status == 0 || status == 3) {
    // DIIS history beyond the device bookkeeping, or an SP2 budget too small: synchronised driver
    // (FP64 Fock builds throughout)
    if (int rc = setup(d, H, S)) return rc;
    if (int rc = scf_synced(d, r, &status, &energy, &iterations)) return rc;
  } else if (r.mixed) {
    // mixed path: one full FP64 build from the last density; its energy is variational (error second
    // order in the density error) and its eigenvectors are the returned orbitals
    if (int rc = zq_fock_jk_device(ctx, ctx->d_P, ctx->d_H, r.x_exch, ctx->d_F)) return rc;
    hipLaunchKernelGGL(k_energy_part, dim3(kRedBlocks), dim3(256), 0, ctx->stream, ctx->d_P, np, ctx->d_H, ctx->d_F, n,
                       d.part, nullptr);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, ctx->stream, d.part, 1, d.scal);
    double h = 0.0;
    if (int rc = sync_read(ctx, d.scal, &h, 1)) return rc;
    energy = h + VN;
    ++iterations;
    ZQ_TRY(ctx, hipMemcpyAsync(d.W, ctx->d_F, nn * 8, hipMemcpyDeviceToDevice, ctx->stream));
  }

  // outputs: C and e from the last diagonalised F (W), P rebuilt from C, and that F
  if (int rc = to_orthonormal(d, d.W)) return rc;
  if (int rc = diag_orthonormal(d, d.C)) return rc;
  hipLaunchKernelGGL(k_density, dim3(kRedBlocks), dim3(256), 0, ctx->stream, d.C, n, nocc, r.factor, ctx->d_P, np, d.part);
  std::vector<double> hP(size_t(np) * np);
  ZQ_TRY(ctx, hipMemcpyAsync(C, d.C, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(e, d.eig, size_t(n) * 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(hP.data(), ctx->d_P, hP.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(F, d.W, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < n; ++i) std::memcpy(P + size_t(i) * n, hP.data() + size_t(i) * np, n * sizeof(double));
  res->energy = energy;
  res->iterations = iterations;
  if (status != 1) return zq_fail(ctx, ZQ_ERR_NOT_CONVERGED, "SCF did not converge");
  ctx->res_C = d.C, ctx->res_e = d.eig, ctx->res_n = n, ctx->res_set = ctx->set_id, ctx->res_gen = ctx->eri_gen;
  return ZQ_OK;
}

extern "C" int zq_scf_set_sp2_budget(zq_ctx *ctx, int steps) {
  if (!ctx || steps < 1 || steps > kSp2MaxSteps)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_scf_set_sp2_budget: steps outside 1..128");
  ctx->sp2_budget = steps;
  return ZQ_OK;
}

extern "C" int zq_scf_set_sp2_guess_budget(zq_ctx *ctx, int steps) {
  if (!ctx || steps < 1 || steps > kSp2MaxSteps)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_scf_set_sp2_guess_budget: steps outside 1..128");
  ctx->sp2_guess_budget = steps;
  return ZQ_OK;
}

extern "C" int zq_scf_set_diis_error_scale(zq_ctx *ctx, double scale) {
  if (!ctx || !(scale >= 0.0 && scale <= 1.0))
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_scf_set_diis_error_scale: scale outside [0, 1]");
  ctx->diis_err_scale = scale;
  return ZQ_OK;
}

extern "C" int zq_scf_last_diis_resets(const zq_ctx *ctx) { return ctx ? ctx->scf_diis_resets : 0; }

extern "C" int zq_scf_last_replays(const zq_ctx *ctx) { return ctx ? ctx->scf_replays : 0; }

extern "C" int zq_scf_last_driver(const zq_ctx *ctx) { return ctx ? ctx->scf_driver : 0; }

extern "C" int zq_scf_captures(const zq_ctx *ctx) { return ctx ? int(ctx->scf_captures) : 0; }

extern "C" int zq_scf_last_sp2(const zq_ctx *ctx, int *densities, int *folded, int *steps) {
  if (!ctx || !densities || !folded || !steps) return ZQ_ERR_INVALID;
  *densities = ctx->scf_sp2[0], *folded = ctx->scf_sp2[1], *steps = ctx->scf_sp2[2];
  return ZQ_OK;
}
