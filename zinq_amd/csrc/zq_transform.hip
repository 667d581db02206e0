// AO->MO four-index transform of the resident ERIs and the MP2 energy that consumes it.
//
// Reference: ao2mo_pppp (src/integral_transform.zig:37-79) and its occupied/virtual restriction
// ao2mo_oovv (:165-219) compute, as four quarter transforms in physicist order,
//   out[i,j,a,b] = sum C[mu,i] C[lam,j] C[nu,a] C[sig,b] g[mu,lam,nu,sig],  g[mu,lam,nu,sig] = (mu nu|lam sig).
// Here the same chain runs as four GEMM stages on MFMA (FP64: v_mfma_f64_16x16x4f64):
//   Q1  X[K][i][nu]      = sum_mu  C1[mu,i]  (mu nu|K)         for every canonical ket pair K = (lam >= sig)
//   Q2  Y[i][j][sig][nu] = sum_lam C2[lam,j] X[(lam sig)][i][nu]   (X is symmetric in lam, sig)
//   Q3  W[i][j][sig][a]  = sum_nu  Y[i][j][sig][nu] C3[nu,a]
//   Q4  out[i][j][a][b]  = sum_sig W[i][j][sig][a] C4[sig,b]
// Q1 reads the 8x8 block quartets straight from HBM into LDS (no unpacked copy of the ERIs) and uses
// the ket-pair symmetry, so its executed flops are half of the dense chain's first quarter.

#include <algorithm>
#include <cmath>

#include "zq_internal.h"

int zq_fock_jk_device(zq_ctx *ctx, const double *dP, const double *dH, double x_exch, double *dF);

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

// v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and B[l/16][l%16] and receives
// D[4*j + l/16][l%16] in accumulator element j (measured on gfx950; the FP32/BF16 16x16 forms use
// row 4*(l/16) + j instead).
__device__ inline d4 mfma_f64(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline int mfma_f64_row(int lane, int j) { return 4 * j + (lane >> 4); }

// Canonical slot of AO pair (l, s): Jb*64 + local, with the larger index as the "r" of the block pair.
__device__ inline size_t slot_of(int l, int s) {
  const int hi = l > s ? l : s, lo = l > s ? s : l;
  return size_t(zq_bpair(hi >> 3, lo >> 3)) * 64 + (hi & 7) * 8 + (lo & 7);
}

// ---------------------------------------------------------------- Q1
// Workgroup (512 threads) owns ket block pair Jb (64 slots) x AO block bnu (8 nu) x a group of 64 i.
// Loop over mu blocks: stage the 64 x 64 tile T[K][u*8+w] = (mu=bmu*8+u, nu=bnu*8+w | K) into LDS from
// the quartet that stores it, then X[K][i][nu] += C1[mu,i] T via MFMA.  Output tiles: MT (<= 4) tiles of
// 16 i by 32 tiles of 16 (K,nu) columns, split evenly over the 8 waves.
template <int MT>
__global__ __launch_bounds__(512) void zq_q1_f64(const double *__restrict__ eri, const double *__restrict__ C1, int ldc1,
                                                 int m1, int nb, int npad, double *__restrict__ X) {
  constexpr int TPW = MT * 4;  // tiles per wave
  __shared__ double T[64][65];
  const int Jb = blockIdx.x / nb, bnu = blockIdx.x % nb;
  const int i0 = blockIdx.y * 64;
  int br, bs;
  zq_unpair(Jb, br, bs);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  d4 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};

  for (int bmu = 0; bmu < nb; ++bmu) {
    const int bmax = bmu > bnu ? bmu : bnu, bmin = bmu > bnu ? bnu : bmu;
    const bool swp = bmu < bnu;
    const int Ib = zq_bpair(bmax, bmin);
    const size_t base =
        (Ib >= Jb ? zq_blk_index(bmax, bmin, br, bs) : zq_blk_index(br, bs, bmax, bmin)) * size_t(ZQ_QUARTET);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = tid + 512 * j;
      const int K = e >> 6, uw = e & 63, u = uw >> 3, w = uw & 7;
      const int r = K >> 3, s = K & 7;
      double v = 0.0;
      if (!(br == bs && r < s)) {
        int pl = swp ? w : u, ql = swp ? u : w;
        if (bmax == bmin && pl < ql) {
          const int t = pl;
          pl = ql, ql = t;
        }
        const int bl = pl * 8 + ql;
        int idx;
        if (Ib > Jb)
          idx = K * 64 + bl;
        else if (Ib < Jb)
          idx = bl * 64 + K;
        else
          idx = (K < bl ? K : bl) * 64 + (K < bl ? bl : K);
        v = eri[base + idx];
      }
      T[K][uw] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int mu = bmu * 8 + ks * 4 + (lane >> 4);
      const int uu = ks * 4 + (lane >> 4);
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        const int tile = wave * TPW + t, it = tile >> 5, nt = tile & 31;
        const int i = i0 + it * 16 + (lane & 15);
        const double a = (i < m1) ? C1[size_t(mu) * ldc1 + i] : 0.0;
        const int col = nt * 16 + (lane & 15), K = col >> 3, wl = col & 7;
        const double b = T[K][uu * 8 + wl];
        acc[t] = mfma_f64(a, b, acc[t]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = wave * TPW + t, it = tile >> 5, nt = tile & 31;
    const int col = nt * 16 + (lane & 15), K = col >> 3, wl = col & 7;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = i0 + it * 16 + mfma_f64_row(lane, j);
      if (i < m1) X[(size_t(Jb) * 64 + K) * size_t(m1) * npad + size_t(i) * npad + bnu * 8 + wl] = acc[t][j];
    }
  }
}

// ---------------------------------------------------------------- generic batched GEMM (FP64 MFMA)
// C[b][m][n] = sum_k A[b][m][k] B[b][k][n] with arbitrary element strides.  Mode 1 gathers B from X:
// B[lam][col = sig*npad + nu] = X[slot(lam, sig)][b][nu] with X strides (slot: sXs, b: sXb).
struct GemmArgs {
  const double *A;
  long long sAb, sAm, sAk;
  const double *B;
  long long sBb, sBk, sBn;
  double *C;
  long long sCb, sCm, sCn;
  int M, N, K;
  int mode;
  int npad;
  long long sXs, sXb;
};

__global__ __launch_bounds__(256) void zq_gemm_f64(GemmArgs g) {
  __shared__ double As[16][65];
  __shared__ double Bs[16][65];
  const int b = blockIdx.z;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const double *A = g.A + b * g.sAb;
  const double *B = g.B + b * g.sBb;
  d4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
  const bool a_kfast = (g.sAk == 1);
  const bool b_nfast = (g.mode == 1) || (g.sBn == 1);
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j;
      int mm, kk;
      if (a_kfast)
        mm = e >> 4, kk = e & 15;
      else
        kk = e >> 6, mm = e & 63;
      const int m = m0 + mm, k = k0 + kk;
      As[kk][mm] = (m < g.M && k < g.K) ? A[m * g.sAm + k * g.sAk] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j;
      int nn, kk;
      if (b_nfast)
        kk = e >> 6, nn = e & 63;
      else
        nn = e >> 4, kk = e & 15;
      const int n = n0 + nn, k = k0 + kk;
      double v = 0.0;
      if (n < g.N && k < g.K) {
        if (g.mode == 1) {
          const int sig = n / g.npad, nu = n - sig * g.npad;
          v = B[(long long)slot_of(k, sig) * g.sXs + nu];
        } else {
          v = B[k * g.sBk + n * g.sBn];
        }
      }
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      double a[2], bb[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) a[x] = As[kk][wm + x * 16 + (lane & 15)];
#pragma unroll
      for (int y = 0; y < 2; ++y) bb[y] = Bs[kk][wn + y * 16 + (lane & 15)];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = mfma_f64(a[x], bb[y], acc[x][y]);
    }
  }
  double *C = g.C + b * g.sCb;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int n = n0 + wn + y * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int m = m0 + wm + x * 16 + mfma_f64_row(lane, j);
        if (m < g.M && n < g.N) C[m * g.sCm + n * g.sCn] = acc[x][y][j];
      }
    }
}

// ---------------------------------------------------------------- MP2 energy
// One workgroup per (i,j): sum_ab T[i,j,a,b] (2 T[i,j,a,b] - T[i,j,b,a]) / (e_i + e_j - e_a - e_b)
// in a fixed order; partial[i*o + j].
__global__ __launch_bounds__(256) void zq_mp2_pairs(const double *__restrict__ T, const double *__restrict__ e, int o,
                                                    int v, double *__restrict__ partial) {
  const int ij = blockIdx.x, i = ij / o, j = ij % o;
  const double *t = T + size_t(ij) * v * v;
  const double eij = e[i] + e[j];
  double s = 0.0;
  for (int ab = threadIdx.x; ab < v * v; ab += 256) {
    const int a = ab / v, b = ab % v;
    const double x = t[ab], y = t[b * v + a];
    s += x * (2.0 * x - y) / (eij - e[o + a] - e[o + b]);
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[ij] = red[0];
}

__global__ void zq_sum_serial(const double *__restrict__ x, int n, double *__restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i];
    *out = s;
  }
}

int launch_gemm(zq_ctx *ctx, const GemmArgs &g, int batch, const char *name) {
  dim3 grid(unsigned((g.N + 63) / 64), unsigned((g.M + 63) / 64), unsigned(batch));
  zq_prof_begin(ctx, name);
  hipLaunchKernelGGL(zq_gemm_f64, grid, dim3(256), 0, ctx->stream, g);
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, name);
  return ZQ_OK;
}

}  // namespace

// Runs Q1..Q4 into device memory; *d_out receives out[i][j][a][b] with dims m1 x m2 x m3 x m4.
int zq_transform_device(zq_ctx *ctx, const double *C, int nocc, int block, int precision, double **d_out, int dims[4]) {
  if (!ctx->eri) return zq_fail(ctx, ZQ_ERR_STATE, "no ERIs loaded");
  const int n = ctx->n, np = ctx->npad, nb = ctx->nb, Nb = ctx->Nb;
  if (block != ZQ_BLOCK_FULL && block != ZQ_BLOCK_OVOV) return zq_fail(ctx, ZQ_ERR_INVALID, "bad block");
  if (precision != ZQ_PREC_FP64) return zq_fail(ctx, ZQ_ERR_INVALID, "precision not available");
  if (block == ZQ_BLOCK_OVOV && (nocc <= 0 || nocc >= n)) return zq_fail(ctx, ZQ_ERR_INVALID, "bad nocc");
  const int m1 = block == ZQ_BLOCK_OVOV ? nocc : n, m2 = m1;
  const int off3 = block == ZQ_BLOCK_OVOV ? nocc : 0, m3 = n - off3, m4 = m3;
  dims[0] = m1, dims[1] = m2, dims[2] = m3, dims[3] = m4;

  // coefficient slices, zero rows for padded AOs: C1 (np x m1), C2T (m2 x np), C3 (np x m3), C4 = C3
  std::vector<double> h1(size_t(np) * m1, 0.0), h2t(size_t(m2) * np, 0.0), h3(size_t(np) * m3, 0.0);
  for (int mu = 0; mu < n; ++mu) {
    for (int i = 0; i < m1; ++i) h1[size_t(mu) * m1 + i] = C[size_t(mu) * n + i];
    for (int j = 0; j < m2; ++j) h2t[size_t(j) * np + mu] = C[size_t(mu) * n + j];
    for (int a = 0; a < m3; ++a) h3[size_t(mu) * m3 + a] = C[size_t(mu) * n + off3 + a];
  }
  double *d1 = (double *)zq_scratch(ctx, "C1", h1.size() * 8);
  double *d2t = (double *)zq_scratch(ctx, "C2T", h2t.size() * 8);
  double *d3 = (double *)zq_scratch(ctx, "C3", h3.size() * 8);
  const size_t nX = size_t(Nb) * 64 * m1 * np, nY = size_t(m1) * m2 * np * np, nW = size_t(m1) * m2 * np * m3,
               nO = size_t(m1) * m2 * m3 * m4;
  double *X = (double *)zq_scratch(ctx, "X", nX * 8);
  double *Y = (double *)zq_scratch(ctx, "Y", nY * 8);
  double *W = (double *)zq_scratch(ctx, "W", nW * 8);
  double *O = (double *)zq_scratch(ctx, "OUT", nO * 8);
  if (!d1 || !d2t || !d3 || !X || !Y || !W || !O) return zq_fail(ctx, ZQ_ERR_OOM, "transform scratch");
  ZQ_TRY(ctx, hipMemcpyAsync(d1, h1.data(), h1.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(d2t, h2t.data(), h2t.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(d3, h3.data(), h3.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));  // the host staging vectors go out of scope below

  // Q1
  {
    const int groups = (m1 + 63) / 64;
    const int mt = std::min(4, (std::min(m1, 64) + 15) / 16);
    dim3 grid(unsigned(Nb * nb), unsigned(groups));
    zq_prof_begin(ctx, "ao2mo_q1");
    switch (mt) {
      case 1: hipLaunchKernelGGL(zq_q1_f64<1>, grid, dim3(512), 0, ctx->stream, ctx->eri, d1, m1, m1, nb, np, X); break;
      case 2: hipLaunchKernelGGL(zq_q1_f64<2>, grid, dim3(512), 0, ctx->stream, ctx->eri, d1, m1, m1, nb, np, X); break;
      case 3: hipLaunchKernelGGL(zq_q1_f64<3>, grid, dim3(512), 0, ctx->stream, ctx->eri, d1, m1, m1, nb, np, X); break;
      default: hipLaunchKernelGGL(zq_q1_f64<4>, grid, dim3(512), 0, ctx->stream, ctx->eri, d1, m1, m1, nb, np, X); break;
    }
    zq_prof_end(ctx);
    ZQ_LAUNCH_CHECK(ctx, "zq_q1_f64");
  }
  // Q2: batch over i; A = C2T (m2 x np), B gathered from X, out Y[i][j][sig][nu]
  {
    GemmArgs g{};
    g.A = d2t, g.sAb = 0, g.sAm = np, g.sAk = 1;
    g.B = X, g.sBb = np, g.mode = 1, g.npad = np, g.sXs = (long long)m1 * np;
    g.C = Y, g.sCb = (long long)m2 * np * np, g.sCm = (long long)np * np, g.sCn = 1;
    g.M = m2, g.N = np * np, g.K = np;
    if (int rc = launch_gemm(ctx, g, m1, "ao2mo_q2")) return rc;
  }
  // Q3: batch over (i,j); A = Y_ij [sig][nu], B = C3 [nu][a], out W[i][j][sig][a]
  {
    GemmArgs g{};
    g.A = Y, g.sAb = (long long)np * np, g.sAm = np, g.sAk = 1;
    g.B = d3, g.sBb = 0, g.sBk = m3, g.sBn = 1;
    g.C = W, g.sCb = (long long)np * m3, g.sCm = m3, g.sCn = 1;
    g.M = np, g.N = m3, g.K = np;
    if (int rc = launch_gemm(ctx, g, m1 * m2, "ao2mo_q3")) return rc;
  }
  // Q4: batch over (i,j); A[a][sig] = W[sig][a], B = C4 [sig][b], out[i][j][a][b]
  {
    GemmArgs g{};
    g.A = W, g.sAb = (long long)np * m3, g.sAm = 1, g.sAk = m3;
    g.B = d3, g.sBb = 0, g.sBk = m4, g.sBn = 1;
    g.C = O, g.sCb = (long long)m3 * m4, g.sCm = m4, g.sCn = 1;
    g.M = m3, g.N = m4, g.K = np;
    if (int rc = launch_gemm(ctx, g, m1 * m2, "ao2mo_q4")) return rc;
  }
  *d_out = O;
  return ZQ_OK;
}

extern "C" int zq_ao2mo(zq_ctx *ctx, const double *C, size_t nocc, int block, int precision, double *out) {
  if (!ctx || !C || !out) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_ao2mo: null argument");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  double *d = nullptr;
  int dims[4];
  if (int rc = zq_transform_device(ctx, C, int(nocc), block, precision, &d, dims)) return rc;
  const size_t cnt = size_t(dims[0]) * dims[1] * dims[2] * dims[3];
  ZQ_TRY(ctx, hipMemcpyAsync(out, d, cnt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ZQ_OK;
}

int zq_mp2_device(zq_ctx *ctx, const double *C, const double *e, int nocc, int precision, double *e_corr) {
  double *T = nullptr;
  int dims[4];
  if (int rc = zq_transform_device(ctx, C, nocc, ZQ_BLOCK_OVOV, precision, &T, dims)) return rc;
  const int o = dims[0], v = dims[2];
  double *de = (double *)zq_scratch(ctx, "mp2_e", size_t(ctx->n) * 8);
  double *part = (double *)zq_scratch(ctx, "mp2_part", size_t(o) * o * 8 + 8);
  if (!de || !part) return zq_fail(ctx, ZQ_ERR_OOM, "mp2 scratch");
  ZQ_TRY(ctx, hipMemcpyAsync(de, e, size_t(ctx->n) * 8, hipMemcpyHostToDevice, ctx->stream));
  zq_prof_begin(ctx, "mp2_energy");
  hipLaunchKernelGGL(zq_mp2_pairs, dim3(unsigned(o * o)), dim3(256), 0, ctx->stream, T, de, o, v, part);
  hipLaunchKernelGGL(zq_sum_serial, dim3(1), dim3(64), 0, ctx->stream, part, o * o, part + size_t(o) * o);
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, "zq_mp2_pairs");
  ZQ_TRY(ctx, hipMemcpyAsync(e_corr, part + size_t(o) * o, 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ZQ_OK;
}

extern "C" int zq_mp2(zq_ctx *ctx, const double *C, const double *e, size_t nocc, int precision, double *e_corr) {
  if (!ctx || !C || !e || !e_corr) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_mp2: null argument");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  return zq_mp2_device(ctx, C, e, int(nocc), precision, e_corr);
}
