// Symmetric eigen-decomposition of the SCF's small dense matrices (n <= 192), the work of eigh /
// geigh in src/linear_algebra.zig:45-71, :184-199 (LAPACKE_dsyevd / dsygvd) on the GPU.
//
// rocSOLVER's dsyevd spends most of its time at these sizes launching the many small kernels of its
// tridiagonal reduction (about 2.4 ms of 3.9 ms at n = 192).  Here the reduction is one workgroup:
//
// k_tridiag: 512 threads hold the n x n matrix in registers (thread (tr, tc) = (tid / 8, tid % 8) owns
//   rows tr + 64 a and columns tc + 8 b), two waves per SIMD, so the per-step scalar work is issued
//   twice per SIMD rather than for every one of many waves.  Householder reduction of the lower
//   triangle as LAPACK dsytd2 (A = Q T Q^T, Q = H_0 ... H_(n-3)), step k:
//     row k -> LDS; alpha = A[k+1][k], beta = -sign(alpha) |A[k+1:][k]|, tau = (beta - alpha) / beta,
//     v = (1, A[k+2:][k] / (alpha - beta));  p = tau A v;  w = p - tau/2 (p.v) v;  A -= v w^T + w v^T,
//   two barriers a step; every wave computes the step's scalars itself, in the same order, so all agree
//   bitwise; column blocks wholly left of k are skipped.  The reflectors stay in LDS until the end,
//   when they go to global memory.
// k_tridiag_q: Q^T = H_(n-3) ... H_0 accumulated in registers (same layout) by right multiplications
//   M <- M H_k, whose row sums (quad sums) need no barrier.  It runs on a second stream while the
//   tridiagonal problem is solved, which needs only d and e.
// The tridiagonal problem T = Z diag(w) Z^T is rocSOLVER dstedc's divide and conquer, and the
// eigenvectors Q Z are one MFMA product once both streams are done.

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/zinq_scf.h"
#include "zq_internal.h"

namespace {

constexpr int kTridiagMax = 192;  // 16 * the largest RB instantiated

// DPP lane moves within a 16-lane row; a source outside the row reads 0.
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over the 16 lanes of a DPP row, complete in the row's last lane (tc == 15).
__device__ inline double row_sum_last(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm 1 0 3 2
  v += dpp_move<0x4E>(v);   // quad_perm 2 3 0 1: every lane holds its quad's sum
  v += dpp_move<0x114>(v);  // row_shr 4
  v += dpp_move<0x118>(v);  // row_shr 8: lanes 12..15 hold the row's sum
  return v;
}

// Offset of reflector k in the packed store: entries k+1 .. n-1.
__device__ inline int vofs(int k, int n) { return k * (n - 1) - k * (k - 1) / 2; }

// A (n x n symmetric, ld lda) -> d (n), e (n - 1) and the reflectors R (refl_words(n): tau, scal, the
// packed row tails) of A = Q T Q^T, n <= 16 RB.
// One workgroup of 16 TR threads per matrix of a strided batch.  Thread (tr, tc) = (tid / 16, tid % 16)
// owns rows tr + TR a (a < RA) and columns tc + 16 b (b < RB), at most 72 doubles (TR = 32 keeps two
// waves per SIMD, so they fit the 256 registers a lane then has); a DPP row of 16 lanes holds one
// thread row, so a row sum is four DPP steps.  Row k (= column k, A symmetric) reaches
// LDS from the 16 lanes that own it, together with |A[k+2:][k]|^2, which they sum on the way.  p.v is
// summed per wave into LDS before the step's second barrier and then added up in a fixed order, so no
// step waits on a 64-lane shuffle reduction.  Row vectors in LDS are stored column-permuted,
// [c % 16][c / 16] with a padded stride, so a lane reads its own columns contiguously and the 16
// column groups fall in different banks.  The reflectors (raw row tails and 1 / (alpha - beta)), d, e
// and tau stay in LDS (at most 156 KB for n = 192) until the loop has ended, so no barrier waits for a
// global store.
template <int RB, int TR>
__global__ __launch_bounds__(16 * TR) void k_tridiag(const double *__restrict__ A, int lda, long long sA, int n,
                                                  double *__restrict__ d, double *__restrict__ e, long long sDE,
                                                  double *__restrict__ R, long long sR) {
  constexpr int NT = 16 * TR, NM = 16 * RB, RA = (NM + TR - 1) / TR, CS = RB + 1;
  __shared__ double cbp[2][16 * CS];  // row k, permuted
  __shared__ double pvp[16 * CS];     // p = tau A v, permuted
  __shared__ double xn[2], pdot[2][TR / 4];
  __shared__ double taus[NM], scals[NM], dl[NM], el[NM];
  __shared__ double vl[(NM - 2) * (NM + 1) / 2];
  A += blockIdx.x * sA;
  d += blockIdx.x * sDE;
  e += blockIdx.x * sDE;
  R += blockIdx.x * sR;
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15, wv = tid >> 6;
  auto P = [](int j) { return (j & 15) * CS + (j >> 4); };
  double a[RA][RB];
#pragma unroll
  for (int ai = 0; ai < RA; ++ai)
#pragma unroll
    for (int bj = 0; bj < RB; ++bj) {
      const int r = tr + TR * ai, c = tc + 16 * bj;
      a[ai][bj] = (r < n && c < n) ? A[size_t(r) * lda + c] : 0.0;
    }
  // p stays 0 on the padding columns (no row >= n writes it), so the update leaves them 0
  for (int t = tid; t < 16 * CS; t += NT) pvp[t] = 0.0;
  // row k of the registers -> dst (permuted) by the 16 lanes that own it, and |A[k][k+2:]|^2 -> *xsq.
  // (the row block is picked with 0/1 weights: branches per block get merged into an indexed access,
  // which moves the register array to scratch memory)
  auto gather_row = [&](int k, double *dst, double *xsq) {
    if (tr == k % TR) {
      const int ak = k / TR;
      double s2 = 0.0;
#pragma unroll
      for (int bj = 0; bj < RB; ++bj) {
        double x = 0.0;
#pragma unroll
        for (int ai = 0; ai < RA; ++ai) x = fma(ai == ak ? 1.0 : 0.0, a[ai][bj], x);
        dst[tc * CS + bj] = x;
        s2 += tc + 16 * bj >= k + 2 ? x * x : 0.0;
      }
      s2 = row_sum_last(s2);
      if (tc == 15 && xsq) *xsq = s2;
    }
  };

  for (int k = 0; k + 2 < n; ++k) {
    const int par = k & 1;
    double *cb = cbp[par];
    gather_row(k, cb, &xn[par]);
    __syncthreads();
    const double alpha = cb[P(k + 1)], xnorm2 = xn[par];
    double beta = alpha, tau = 0.0, scal = 0.0;
    if (xnorm2 != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    if (tid == 0) {
      dl[k] = cb[P(k)];
      el[k] = beta;
      taus[k] = tau;
      scals[k] = scal;
    }
    // raw tail of row k for the back-accumulation: v_k[j] = scal * tail[j] for j > k + 1, v_k[k+1] = 1
    for (int j = k + 1 + tid; j < n; j += NT) vl[vofs(k, n) + j - k - 1] = cb[P(j)];
    auto vval = [&](int i) { return i == k + 1 ? 1.0 : (i > k + 1 && i < n ? cb[P(i)] * scal : 0.0); };
    // v on this lane's columns (0 at and left of k), read from the row in LDS at each use: registers
    // hold A
    auto vcol = [&](int bj) {
      const int j = tc + 16 * bj;
      const double x = cb[tc * CS + bj] * scal;
      return j > k + 1 ? x : (j == k + 1 ? 1.0 : 0.0);
    };
    // p = tau A v on this lane's rows; the row's last lane stores p_i and adds p_i v_i to its partial
    double pt[RA];
#pragma unroll
    for (int ai = 0; ai < RA; ++ai) pt[ai] = 0.0;
#pragma unroll
    for (int bj = 0; bj < RB; ++bj) {
      const double v = vcol(bj);
#pragma unroll
      for (int ai = 0; ai < RA; ++ai) pt[ai] += a[ai][bj] * v;
      if (bj % 4 == 3) __builtin_amdgcn_sched_barrier(0);  // bounds the loads in flight (registers hold A)
    }
    double pd = 0.0;
#pragma unroll
    for (int ai = 0; ai < RA; ++ai) {
      const double t = tau * row_sum_last(pt[ai]);
      const int r = tr + TR * ai;
      if (tc == 15 && r > k && r < n) {
        pvp[P(r)] = t;
        pd += t * vval(r);
      }
    }
    // per-wave sum of p.v over the wave's 4 row-last lanes (15, 31, 47, 63)
    pd = pd + __shfl_xor(pd, 16, 64);
    pd = pd + __shfl_xor(pd, 32, 64);
    if ((tid & 63) == 63) pdot[par][wv] = pd;
    __syncthreads();
    double dot = 0.0;
#pragma unroll
    for (int q = 0; q < TR / 4; ++q) dot += pdot[par][q];
    const double K = -0.5 * tau * dot;
    double vr[RA], wr[RA];
#pragma unroll
    for (int ai = 0; ai < RA; ++ai) {
      const int r = tr + TR * ai;
      vr[ai] = vval(r);
      wr[ai] = (r > k && r < n) ? pvp[P(r)] + K * vr[ai] : 0.0;
    }
    // A -= v w^T + w v^T on the trailing block (w = 0 at and left of k)
#pragma unroll
    for (int bj = 0; bj < RB; ++bj) {
      const double v = vcol(bj), w = tc + 16 * bj > k ? pvp[tc * CS + bj] + K * v : 0.0;
#pragma unroll
      for (int ai = 0; ai < RA; ++ai) a[ai][bj] -= vr[ai] * w + wr[ai] * v;
      if (bj % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    }
  }
  // the last 2 x 2 block (or the 1 x 1 matrix)
  __syncthreads();
  gather_row(n >= 2 ? n - 2 : 0, cbp[0], nullptr);
  if (n >= 2) gather_row(n - 1, cbp[1], nullptr);
  __syncthreads();
  if (tid == 0) {
    if (n >= 2) {
      dl[n - 2] = cbp[0][P(n - 2)];
      el[n - 2] = cbp[1][P(n - 2)];
      dl[n - 1] = cbp[1][P(n - 1)];
    } else {
      dl[0] = cbp[0][P(0)];
    }
  }
  __syncthreads();  // reflectors, taus, d and e complete
  for (int t = tid; t < n; t += NT) d[t] = dl[t];
  for (int t = tid; t + 1 < n; t += NT) e[t] = el[t];
  // the reflectors for k_tridiag_q: tau_k and scal_k (k < n - 2), then the packed row tails
  const int nk = n > 2 ? n - 2 : 0, nv = n > 2 ? vofs(n - 2, n) : 0;
  for (int t = tid; t < nk; t += NT) R[t] = taus[t], R[n + t] = scals[t];
  for (int t = tid; t < nv; t += NT) R[2 * n + t] = vl[t];
}

// Q from the reflectors k_tridiag left in R (one workgroup per matrix of the batch, the same thread
// layout): M = Q^T = H_(n-3) ... H_0, from M = I by M <- M H_k = M - tau (M v) v^T for k = n-3 .. 0,
// whose row sums (quad sums) need no barrier; then Q[c][r] = M[r][c].  It runs beside the tridiagonal
// eigensolver, which needs only d and e.
template <int RB, int TR>
__global__ __launch_bounds__(16 * TR) void k_tridiag_q(const double *__restrict__ R, long long sR, int n,
                                                    double *__restrict__ Q, long long sQ) {
  constexpr int NT = 16 * TR, NM = 16 * RB, RA = (NM + TR - 1) / TR;
  __shared__ double taus[NM], scals[NM];
  __shared__ double vl[(NM - 2) * (NM + 1) / 2];
  R += blockIdx.x * sR;
  Q += blockIdx.x * sQ;
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  const int nk = n > 2 ? n - 2 : 0, nv = n > 2 ? vofs(n - 2, n) : 0;
  for (int t = tid; t < nk; t += NT) taus[t] = R[t], scals[t] = R[n + t];
  for (int t = tid; t < nv; t += NT) vl[t] = R[2 * n + t];
  __syncthreads();
  double a[RA][RB];
#pragma unroll
  for (int ai = 0; ai < RA; ++ai)
#pragma unroll
    for (int bj = 0; bj < RB; ++bj) {
      const int r = tr + TR * ai, c = tc + 16 * bj;
      a[ai][bj] = (r == c && r < n) ? 1.0 : 0.0;
    }
  for (int k = n - 3; k >= 0; --k) {
    const double tau = taus[k], scal = scals[k];
    const double *tail = vl + vofs(k, n) - k - 1;  // tail[j], j > k
    auto vcol = [&](int bj) {
      const int j = tc + 16 * bj, jc = min(max(j, k + 1), n - 1);  // reads stay inside reflector k
      const double x = tail[jc] * scal;
      return (j > k + 1 && j < n) ? x : (j == k + 1 ? 1.0 : 0.0);
    };
    // rows <= k of M are still those of I and v vanishes there, so their u is 0
    double u[RA];
#pragma unroll
    for (int ai = 0; ai < RA; ++ai) u[ai] = 0.0;
#pragma unroll
    for (int bj = 0; bj < RB; ++bj) {
      const double v = vcol(bj);
#pragma unroll
      for (int ai = 0; ai < RA; ++ai) u[ai] += a[ai][bj] * v;
      if (bj % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int ai = 0; ai < RA; ++ai) u[ai] = tau * __shfl(row_sum_last(u[ai]), (tid & ~15) | 15, 64);
#pragma unroll
    for (int bj = 0; bj < RB; ++bj) {
      const double v = vcol(bj);
#pragma unroll
      for (int ai = 0; ai < RA; ++ai) a[ai][bj] -= u[ai] * v;
      if (bj % 4 == 3) __builtin_amdgcn_sched_barrier(0);
    }
  }
  // Q[c][r] = M[r][c]
#pragma unroll
  for (int ai = 0; ai < RA; ++ai)
#pragma unroll
    for (int bj = 0; bj < RB; ++bj) {
      const int r = tr + TR * ai, c = tc + 16 * bj;
      if (r < n && c < n) Q[size_t(c) * n + r] = a[ai][bj];
    }
}

// Words of the reflector store of one n x n matrix: tau and scal per step, the packed row tails.
inline size_t refl_words(int n) { return 2 * size_t(n) + (n > 2 ? size_t(n - 2) * (n + 1) / 2 : 0); }

// The reduction on s; Q on s2, which waits for it, while s goes on (the caller joins s2 back).
template <int RB, int TR>
hipError_t launch_tridiag(hipStream_t s, hipStream_t s2, hipEvent_t fork, int count, const double *A, int lda,
                          long long sA, int n, double *d, double *e, long long sDE, double *R, double *Q,
                          long long sQ) {
  const long long sR = (long long)refl_words(n);
  hipLaunchKernelGGL((k_tridiag<RB, TR>), dim3(count), dim3(16 * TR), 0, s, A, lda, sA, n, d, e, sDE, R, sR);
  if (hipError_t err = hipEventRecord(fork, s)) return err;
  if (hipError_t err = hipStreamWaitEvent(s2, fork, 0)) return err;
  hipLaunchKernelGGL((k_tridiag_q<RB, TR>), dim3(count), dim3(16 * TR), 0, s2, R, sR, n, Q, sQ);
  return hipGetLastError();
}

}  // namespace

// The context's rocBLAS handle, on its stream, with atomics off so that every product is reproducible.
void *zq_blas(zq_ctx *ctx) {
  if (!ctx->solver) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    rocblas_set_stream(h, ctx->stream);
    rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed);
    ctx->solver = h;
  }
  return ctx->solver;
}

namespace {

rocblas_handle eig_handle(zq_ctx *ctx) { return (rocblas_handle)zq_blas(ctx); }

// The stream the Q accumulations of every context on a device share, created on first use and kept
// until the process ends: a stream per context would add a hardware queue per context, and streams
// beyond the hardware queues share them, ordering unrelated work.  Two contexts' accumulations may
// queue behind each other here; each still waits only for its own reduction.
hipError_t device_eig_stream(int device, hipStream_t *out) {
  static std::mutex mu;
  static std::map<int, hipStream_t> streams;
  std::lock_guard<std::mutex> lock(mu);
  hipStream_t &st = streams[device];
  if (!st)
    if (hipError_t err = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) {
      st = nullptr;
      return err;
    }
  *out = st;
  return hipSuccess;
}

}  // namespace

int zq_syev_rows(zq_ctx *ctx, double *A, int n, double *w, double *offd, int *info) {
  if (n < 1) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_syev_rows: n < 1");
  rocblas_handle h = eig_handle(ctx);
  if (!h) return zq_fail(ctx, ZQ_ERR_LAPACK, "rocblas handle");
  zq_prof_begin(ctx, "scf_eig");
  if (n > kTridiagMax) {
    const rocblas_status st = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, A, n, w, offd, info);
    zq_prof_end(ctx);
    if (st != rocblas_status_success) return zq_fail(ctx, ZQ_ERR_LAPACK, "rocsolver_dsyevd");
    return ZQ_OK;
  }
  const size_t nn = size_t(n) * n;
  auto *Qm = (double *)zq_scratch(ctx, "eig_Q", nn * sizeof(double));
  auto *Z = (double *)zq_scratch(ctx, "eig_Z", nn * sizeof(double));
  // no captured SCF iteration holds the reflector store, so growing it keeps the capture cache
  auto *R = (double *)zq_scratch(ctx, "eig_refl", refl_words(n) * sizeof(double), nullptr, false);
  if (!Qm || !Z || !R) {
    zq_prof_end(ctx);
    return zq_fail(ctx, ZQ_ERR_OOM, "eigen-solver buffers");
  }
  hipStream_t s2 = nullptr;
  ZQ_TRY(ctx, device_eig_stream(ctx->device, &s2));
  for (hipEvent_t &ev : ctx->eig_events)
    if (!ev) ZQ_TRY(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  hipStream_t s = ctx->stream;
  hipEvent_t fork = ctx->eig_events[0], join = ctx->eig_events[1];
  hipError_t le;
  if (n <= 64)
    le = launch_tridiag<4, 64>(s, s2, fork, 1, A, n, 0, n, w, offd, 0, R, Qm, 0);
  else if (n <= 128)
    le = launch_tridiag<8, 64>(s, s2, fork, 1, A, n, 0, n, w, offd, 0, R, Qm, 0);
  else
    le = launch_tridiag<12, 32>(s, s2, fork, 1, A, n, 0, n, w, offd, 0, R, Qm, 0);
  if (le != hipSuccess) {
    zq_prof_end(ctx);
    return zq_check(ctx, le, "k_tridiag");
  }
  // T = Z diag(w) Z^T: column k of the column-major Z (= row k of Z as stored) is eigenvector k; Q is
  // accumulated on s2 meanwhile
  const rocblas_status st = rocsolver_dstedc(h, rocblas_evect_tridiagonal, n, w, offd, Z, n, info);
  // s waits for Q even when the solver failed, so no later work on s overtakes it
  const hipError_t je = hipEventRecord(join, s2);
  const hipError_t we = je == hipSuccess ? hipStreamWaitEvent(s, join, 0) : je;
  zq_prof_end(ctx);  // before the product, which times itself (profile regions do not nest)
  if (we != hipSuccess) return zq_check(ctx, we, "eigen-solver join");
  if (st != rocblas_status_success) return zq_fail(ctx, ZQ_ERR_LAPACK, "rocsolver_dstedc");
  // row k of A = (Q z_k)^T: A[k][p] = sum_q Z[k][q] Q[p][q]
  return zq_gemm_nt(ctx, Z, n, Qm, n, A, n, n, n, n, 1.0, "scf_eig");
}

extern "C" int zq_eigh(zq_ctx *ctx, size_t n, const double *A, double *W, double *U) {
  ZqGate _gate;
  if (!ctx || !A || !W || !U || n < 1) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_eigh: null argument or n = 0");
  if (n > 8192) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_eigh: n too large");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  const size_t nn = n * n;
  auto *dA = (double *)zq_scratch(ctx, "eigh_A", nn * sizeof(double));
  auto *dw = (double *)zq_scratch(ctx, "eigh_w", 2 * n * sizeof(double));
  auto *dinfo = (int *)zq_scratch(ctx, "eigh_info", sizeof(int));
  if (!dA || !dw || !dinfo) return zq_fail(ctx, ZQ_ERR_OOM, "zq_eigh buffers");
  ZQ_TRY(ctx, hipMemcpyAsync(dA, A, nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  if (int rc = zq_syev_rows(ctx, dA, int(n), dw, dw + n, dinfo)) return rc;
  std::vector<double> V(nn);
  int info = 0;
  ZQ_TRY(ctx, hipMemcpyAsync(V.data(), dA, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(W, dw, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (info != 0) return zq_fail(ctx, ZQ_ERR_LAPACK, "zq_eigh: tridiagonal eigensolver did not converge");
  for (size_t i = 0; i < n; ++i)  // eigenvector k is row k of V and column k of U
    for (size_t k = 0; k < n; ++k) U[i * n + k] = V[k * n + i];
  return ZQ_OK;
}
