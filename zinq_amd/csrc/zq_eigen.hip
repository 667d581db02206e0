// Symmetric eigen-decomposition of the SCF's small dense matrices (n <= 192), the work of eigh /
// geigh in src/linear_algebra.zig:45-71, :184-199 (LAPACKE_dsyevd / dsygvd) on the GPU.
//
// rocSOLVER's dsyevd spends most of its time at these sizes launching the many small kernels of its
// tridiagonal reduction (about 2.4 ms of 3.9 ms at n = 192).  Here the reduction is one workgroup:
//
// k_tridiag: 1024 threads hold the n x n matrix in registers, thread (tr, tc) = (tid / 32, tid % 32)
//   owning rows tr + 32 a and columns tc + 32 b (a, b < R = ceil(n / 32)).  A wave holds two thread
//   rows, so every row reduction is a 32-lane shuffle reduction inside a half-wave.  Householder
//   reduction of the lower triangle as LAPACK dsytd2 (A = Q T Q^T, Q = H_0 ... H_(n-3)), step k:
//     column k -> LDS; alpha = A[k+1][k], beta = -sign(alpha) |A[k+1:][k]|, tau = (beta - alpha) / beta,
//     v = (1, A[k+2:][k] / (alpha - beta));  p = tau A v;  w = p - tau/2 (p.v) v;  A -= v w^T + w v^T,
//   two barriers a step; every wave computes the step's scalars itself, in the same order, so all agree
//   bitwise.  The reflectors go to global memory, and Q^T = H_(n-3) ... H_0 is then accumulated in
//   the same registers by right multiplications M <- M H_k, whose row sums need no barrier.
// The tridiagonal problem T = Z diag(w) Z^T is rocSOLVER dstedc's divide and conquer, and the
// eigenvectors Q Z are one MFMA product.

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <vector>

#include "../../include/zinq_scf.h"
#include "zq_internal.h"

namespace {

constexpr int kMaxR = 6;  // n <= 32 * kMaxR: the matrix and the step's vectors fit 128 VGPRs a lane

// row[q] for a run-time q, as a sum with 0/1 weights: a select between elements would be folded into
// an indexed load, which moves the whole register array to scratch memory.
template <int R>
__device__ inline double pick(const double (&row)[R], int q) {
  double v = 0.0;
#pragma unroll
  for (int t = 0; t < R; ++t) v = fma(t == q ? 1.0 : 0.0, row[t], v);
  return v;
}

template <int R>
__device__ inline double pick2(const double (&m)[R][R], int ri, int ci) {
  double v = 0.0;
#pragma unroll
  for (int t = 0; t < R; ++t) v = fma(t == ri ? 1.0 : 0.0, pick<R>(m[t], ci), v);
  return v;
}

__device__ inline double half_sum(double v) {  // sum over the 32 lanes of this half-wave
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline double wave_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// A (n x n symmetric, ld lda) -> d (n), e (n - 1), Q (n x n row-major, ld n: A = Q T Q^T), with Vh
// (n x n) holding the reflectors.  One workgroup per matrix of a strided batch.
template <int R>
__global__ __launch_bounds__(1024) void k_tridiag(const double *__restrict__ A, int lda, long long sA, int n,
                                                  double *__restrict__ d, double *__restrict__ e, long long sDE,
                                                  double *__restrict__ Q, long long sQ, double *__restrict__ Vh) {
  __shared__ double col[2][32 * R];
  __shared__ double pv[32 * R];
  __shared__ double taus[32 * R];
  A += blockIdx.x * sA;
  d += blockIdx.x * sDE;
  e += blockIdx.x * sDE;
  Q += blockIdx.x * sQ;
  Vh += blockIdx.x * sQ;
  const int tid = threadIdx.x, tr = tid >> 5, tc = tid & 31, lane = tid & 63;
  double a[R][R];
#pragma unroll
  for (int ai = 0; ai < R; ++ai)
#pragma unroll
    for (int bj = 0; bj < R; ++bj) {
      const int r = tr + 32 * ai, c = tc + 32 * bj;
      a[ai][bj] = (r < n && c < n) ? A[size_t(r) * lda + c] : 0.0;
    }

  for (int k = 0; k + 2 < n; ++k) {
    double *cb = col[k & 1];
    if (tc == (k & 31)) {
#pragma unroll
      for (int ai = 0; ai < R; ++ai) cb[tr + 32 * ai] = pick<R>(a[ai], k >> 5);
    }
    __syncthreads();
    const double alpha = cb[k + 1];
    double s = 0.0;
    for (int i = k + 2 + lane; i < n; i += 64) s += cb[i] * cb[i];
    const double xnorm2 = wave_sum64(s);
    double beta = alpha, tau = 0.0, scal = 0.0;
    if (xnorm2 != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    auto vval = [&](int i) { return i == k + 1 ? 1.0 : (i > k + 1 && i < n ? cb[i] * scal : 0.0); };
    if (tid == 0) {
      d[k] = cb[k];
      e[k] = beta;
      taus[k] = tau;
    }
    if (tid < n) Vh[size_t(k) * n + tid] = vval(tid);
    double vc[R];
#pragma unroll
    for (int bj = 0; bj < R; ++bj) vc[bj] = vval(tc + 32 * bj);
    // p = tau A v on the rows below k (rows above are not used)
#pragma unroll
    for (int ai = 0; ai < R; ++ai) {
      if (32 * ai + 31 <= k) continue;
      double t = 0.0;
#pragma unroll
      for (int bj = 0; bj < R; ++bj) t += a[ai][bj] * vc[bj];
      t = half_sum(t);
      if (tc == 0) pv[tr + 32 * ai] = tau * t;
    }
    __syncthreads();
    s = 0.0;
    for (int i = k + 1 + lane; i < n; i += 64) s += pv[i] * vval(i);
    const double K = -0.5 * tau * wave_sum64(s);
    double wc[R];
#pragma unroll
    for (int bj = 0; bj < R; ++bj) {
      const int j = tc + 32 * bj;
      wc[bj] = (j > k && j < n) ? pv[j] + K * vc[bj] : 0.0;
    }
#pragma unroll
    for (int ai = 0; ai < R; ++ai) {
      if (32 * ai + 31 <= k) continue;
      const int i = tr + 32 * ai;
      const double vr = vval(i), wr = (i > k && i < n) ? pv[i] + K * vr : 0.0;
#pragma unroll
      for (int bj = 0; bj < R; ++bj) a[ai][bj] -= vr * wc[bj] + wr * vc[bj];
    }
  }
  // the last 2 x 2 block
  __syncthreads();
  if (n >= 2) {
    if (tr == ((n - 2) & 31) && tc == ((n - 2) & 31)) d[n - 2] = pick2<R>(a, (n - 2) >> 5, (n - 2) >> 5);
    if (tr == ((n - 1) & 31) && tc == ((n - 2) & 31)) e[n - 2] = pick2<R>(a, (n - 1) >> 5, (n - 2) >> 5);
  }
  if (tr == ((n - 1) & 31) && tc == ((n - 1) & 31)) d[n - 1] = pick2<R>(a, (n - 1) >> 5, (n - 1) >> 5);
  __syncthreads();  // the reflectors in Vh and taus are complete

  // M = Q^T = H_(n-3) ... H_0, from M = I by M <- M H_k = M - tau (M v) v^T for k = n-3 .. 0
#pragma unroll
  for (int ai = 0; ai < R; ++ai)
#pragma unroll
    for (int bj = 0; bj < R; ++bj) {
      const int r = tr + 32 * ai, c = tc + 32 * bj;
      a[ai][bj] = (r == c && r < n) ? 1.0 : 0.0;
    }
  double vn[R];
  if (n >= 3) {
#pragma unroll
    for (int bj = 0; bj < R; ++bj) {
      const int j = tc + 32 * bj;
      vn[bj] = j < n ? Vh[size_t(n - 3) * n + j] : 0.0;
    }
  }
  for (int k = n - 3; k >= 0; --k) {
    double vc[R];
#pragma unroll
    for (int bj = 0; bj < R; ++bj) vc[bj] = vn[bj];
    if (k > 0) {  // the next reflector's loads overlap this step
#pragma unroll
      for (int bj = 0; bj < R; ++bj) {
        const int j = tc + 32 * bj;
        vn[bj] = j < n ? Vh[size_t(k - 1) * n + j] : 0.0;
      }
    }
    const double tau = taus[k];
#pragma unroll
    for (int ai = 0; ai < R; ++ai) {
      if (32 * ai + 31 <= k) continue;  // rows <= k of M are still those of I, and v vanishes there
      double u = 0.0;
#pragma unroll
      for (int bj = 0; bj < R; ++bj) u += a[ai][bj] * vc[bj];
      u = tau * half_sum(u);
#pragma unroll
      for (int bj = 0; bj < R; ++bj) a[ai][bj] -= u * vc[bj];
    }
  }
  // Q[c][r] = M[r][c]
#pragma unroll
  for (int ai = 0; ai < R; ++ai)
#pragma unroll
    for (int bj = 0; bj < R; ++bj) {
      const int r = tr + 32 * ai, c = tc + 32 * bj;
      if (r < n && c < n) Q[size_t(c) * n + r] = a[ai][bj];
    }
}

template <int R>
void launch_tridiag(hipStream_t s, int count, const double *A, int lda, long long sA, int n, double *d, double *e,
                    long long sDE, double *Q, long long sQ, double *Vh) {
  hipLaunchKernelGGL(k_tridiag<R>, dim3(count), dim3(1024), 0, s, A, lda, sA, n, d, e, sDE, Q, sQ, Vh);
}

rocblas_handle eig_handle(zq_ctx *ctx) {
  if (!ctx->solver) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    rocblas_set_stream(h, ctx->stream);
    ctx->solver = h;
  }
  return (rocblas_handle)ctx->solver;
}

}  // namespace

int zq_syev_rows(zq_ctx *ctx, double *A, int n, double *w, double *offd, int *info) {
  if (n < 1) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_syev_rows: n < 1");
  rocblas_handle h = eig_handle(ctx);
  if (!h) return zq_fail(ctx, ZQ_ERR_LAPACK, "rocblas handle");
  zq_prof_begin(ctx, "scf_eig");
  if (n > 32 * kMaxR) {
    const rocblas_status st = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, A, n, w, offd, info);
    zq_prof_end(ctx);
    if (st != rocblas_status_success) return zq_fail(ctx, ZQ_ERR_LAPACK, "rocsolver_dsyevd");
    return ZQ_OK;
  }
  const size_t nn = size_t(n) * n;
  auto *Qm = (double *)zq_scratch(ctx, "eig_Q", nn * sizeof(double));
  auto *Vh = (double *)zq_scratch(ctx, "eig_V", nn * sizeof(double));
  auto *Z = (double *)zq_scratch(ctx, "eig_Z", nn * sizeof(double));
  if (!Qm || !Vh || !Z) {
    zq_prof_end(ctx);
    return zq_fail(ctx, ZQ_ERR_OOM, "eigen-solver buffers");
  }
  const int R = (n + 31) / 32;
  hipStream_t s = ctx->stream;
  switch (R) {
    case 1: launch_tridiag<1>(s, 1, A, n, 0, n, w, offd, 0, Qm, 0, Vh); break;
    case 2: launch_tridiag<2>(s, 1, A, n, 0, n, w, offd, 0, Qm, 0, Vh); break;
    case 3: launch_tridiag<3>(s, 1, A, n, 0, n, w, offd, 0, Qm, 0, Vh); break;
    case 4: launch_tridiag<4>(s, 1, A, n, 0, n, w, offd, 0, Qm, 0, Vh); break;
    case 5: launch_tridiag<5>(s, 1, A, n, 0, n, w, offd, 0, Qm, 0, Vh); break;
    default: launch_tridiag<6>(s, 1, A, n, 0, n, w, offd, 0, Qm, 0, Vh); break;
  }
  ZQ_LAUNCH_CHECK(ctx, "k_tridiag");
  // T = Z diag(w) Z^T: column k of the column-major Z (= row k of Z as stored) is eigenvector k
  const rocblas_status st = rocsolver_dstedc(h, rocblas_evect_tridiagonal, n, w, offd, Z, n, info);
  if (st != rocblas_status_success) {
    zq_prof_end(ctx);
    return zq_fail(ctx, ZQ_ERR_LAPACK, "rocsolver_dstedc");
  }
  // row k of A = (Q z_k)^T: A[k][p] = sum_q Z[k][q] Q[p][q]
  const int rc = zq_gemm_nt(ctx, Z, n, Qm, n, A, n, n, n, n, 1.0, "scf_eig");
  zq_prof_end(ctx);
  return rc;
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
