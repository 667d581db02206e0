// zq_ci.hip -- the determinant consumers of the FULL AO->MO transform on the GPU:
//   zq_ci   CI states of computeCiStates (src/configuration_interaction.zig:597-683): transformInts
//           (ao2mo_pp + ao2mo_pppp in the spin-orbital basis, :597-632), the Slater-Condon H_CI over
//           generateDets (:126-183, slater :323-432), its eigenvalues + V_NN;
//   zq_mp   Rayleigh-Schroedinger MPn of mp (src/moller_plesset.zig:340-481) in the determinant space of
//           singles .. order-fold excitations;
//   zq_ao2mo_pp  the one-electron transform ao2mo_pp (src/integral_transform.zig:12-34).
//
// The reference expands every integral to the (2n)^4 spin-orbital tensor (ao2so_pppp) before the
// transform.  Here the FULL transform runs once on the n spatial orbitals and the spin-orbital integral
// is read from it on the fly: spin orbital p = 2P + s (the column order of ao2so_coef,
// src/integral_transform.zig:82-101), <pq|rs> = <PQ|RS> when s_p = s_r and s_q = s_s, else 0 -- the
// same values the reference's transform of the spin-blocked tensor produces (its other terms are
// products with exact zeros).  A generalized (GHF) reference has spin-blocked ERIs loaded and C of
// dimension 2 nbf, so its transform is already in spin orbitals.
//
// Determinants are bit masks of occupied spin orbitals (up to 128).  One thread evaluates one H_CI
// element with the Slater-Condon rules; the excitation, its phase and the orbital sums follow the
// reference's order (occupied orbitals ascending).  H_CI is dense FP64 in HBM, diagonalised with the
// context's eigensolver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "zq_internal.h"

namespace {

constexpr int kWords = 2;            // determinant bit-mask words: up to 128 spin orbitals
constexpr size_t kMaxDets = 16384;   // dense H_CI (2 GiB) and its eigensolve
constexpr size_t kMaxRowDets = size_t(1) << 24;  // order-2 MP: only row 0 of H_CI is formed

struct Det {
  unsigned long long w[kWords];
};

struct SoInts {
  const double *g;  // <PQ|RS>, n^4 (FULL transform, physicist order)
  const double *h;  // one-electron operator in spin orbitals, nsp x nsp
  int n, nsp;
  int so;  // 1: spatial g, spin orbital p = 2P + s; 0: g already in spin orbitals
  __device__ double G(int p, int q, int r, int s) const {
    if (so) {
      if (((p ^ r) | (q ^ s)) & 1) return 0.0;
      p >>= 1, q >>= 1, r >>= 1, s >>= 1;
    }
    return g[((size_t(p) * n + q) * n + r) * n + s];
  }
  __device__ double H(int p, int q) const { return h[size_t(p) * nsp + q]; }
};

__device__ inline bool has(const Det &d, int p) { return (d.w[p >> 6] >> (p & 63)) & 1ull; }

// number of occupied spin orbitals of d below p
__device__ inline int below(const Det &d, int p) {
  int c = 0;
#pragma unroll
  for (int k = 0; k < kWords; ++k) {
    const int lo = k * 64;
    if (p >= lo + 64)
      c += __popcll(d.w[k]);
    else if (p > lo)
      c += __popcll(d.w[k] & ((1ull << (p - lo)) - 1ull));
  }
  return c;
}

// lowest set bit of x (nonzero) across the words, cleared
__device__ inline int pop_lowest(unsigned long long x[kWords]) {
#pragma unroll
  for (int k = 0; k < kWords; ++k)
    if (x[k]) {
      const int b = __ffsll((long long)x[k]) - 1;
      x[k] &= x[k] - 1ull;
      return k * 64 + b;
    }
  return -1;
}

// slater(A, B) of src/configuration_interaction.zig:323-432.
__device__ double slater(const Det &A, const Det &B, int nsp, const SoInts &I) {
  unsigned long long a[kWords], b[kWords];
  int ndiff = 0;
#pragma unroll
  for (int k = 0; k < kWords; ++k) {
    a[k] = A.w[k] & ~B.w[k];
    b[k] = B.w[k] & ~A.w[k];
    ndiff += __popcll(a[k]);
  }
  if (ndiff > 2) return 0.0;
  if (ndiff == 0) {
    double sum = 0.0;
    unsigned long long x[kWords];
#pragma unroll
    for (int k = 0; k < kWords; ++k) x[k] = A.w[k];
    for (int i = pop_lowest(x); i >= 0; i = pop_lowest(x)) sum = sum + I.H(i, i);
#pragma unroll
    for (int k = 0; k < kWords; ++k) x[k] = A.w[k];
    for (int i = pop_lowest(x); i >= 0; i = pop_lowest(x)) {
      unsigned long long y[kWords];
#pragma unroll
      for (int k = 0; k < kWords; ++k) y[k] = x[k];  // the occupied orbitals above i
      for (int j = pop_lowest(y); j >= 0; j = pop_lowest(y)) sum = sum + (I.G(i, j, i, j) - I.G(i, j, j, i));
    }
    return sum;
  }
  if (ndiff == 1) {
    const int m = pop_lowest(a), p = pop_lowest(b);
    const int sign = ((below(A, m) + below(B, p)) & 1) ? -1 : 1;
    double term = I.H(m, p);
    unsigned long long x[kWords];
#pragma unroll
    for (int k = 0; k < kWords; ++k) x[k] = A.w[k];
    for (int i = pop_lowest(x); i >= 0; i = pop_lowest(x))
      if (i != m) term = term + (I.G(m, i, p, i) - I.G(m, i, i, p));
    return term * double(sign);
  }
  const int m = pop_lowest(a), nn = pop_lowest(a), p = pop_lowest(b), q = pop_lowest(b);
  const int e = below(A, m) + below(A, nn) - 1 + below(B, p) + below(B, q) - 1;
  return (I.G(m, nn, p, q) - I.G(m, nn, q, p)) * double((e & 1) ? -1 : 1);
}

// H_CI lower triangle and its mirror; rows = ndet (all) or 1 (row 0 only).
__global__ __launch_bounds__(256) void k_hci(const Det *__restrict__ dets, int ndet, int rows, SoInts I,
                                             double *__restrict__ Hci, size_t ld) {
  const int J = blockIdx.x * 16 + (threadIdx.x & 15), Irow = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (Irow >= rows || J >= ndet) return;
  if (rows > 1 && J > Irow) return;
  const double v = slater(dets[Irow], dets[J], I.nsp, I);
  Hci[size_t(Irow) * ld + J] = v;
  if (rows > 1) Hci[size_t(J) * ld + Irow] = v;
}

// One-electron MO integrals in spin orbitals: h_so[p][q] = h[P][Q] for equal spins (so), or h[p][q].
__global__ void k_h_so(const double *__restrict__ h, int n, int nsp, int so, double *__restrict__ hso) {
  const int p = blockIdx.x, q = threadIdx.x;
  if (q >= nsp) return;
  double v;
  if (so)
    v = ((p ^ q) & 1) ? 0.0 : h[size_t(p >> 1) * n + (q >> 1)];
  else
    v = h[size_t(p) * n + q];
  hso[size_t(p) * nsp + q] = v;
}

// H_MS of mp (src/moller_plesset.zig:377-388): diag(e_SO) - sum_{i < nel} (<pi|qi> - <pi|iq>).
__global__ void k_mp_hms(SoInts I, const double *__restrict__ e, int nel, double *__restrict__ hms) {
  const int p = blockIdx.x, q = threadIdx.x;
  if (q >= I.nsp) return;
  double sum = 0.0;
  for (int i = 0; i < nel; ++i) sum = sum + (I.G(p, i, q, i) - I.G(p, i, i, q));
  const double ep = I.so ? e[p >> 1] : e[p];
  hms[size_t(p) * I.nsp + q] = (p == q ? ep : 0.0) - sum;
}

// E0 of every determinant: the sum of its spin-orbital energies in ascending orbital order.
__global__ void k_e0(const Det *__restrict__ dets, int ndet, const double *__restrict__ e, int so,
                     double *__restrict__ E0) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndet) return;
  unsigned long long x[kWords];
#pragma unroll
  for (int k = 0; k < kWords; ++k) x[k] = dets[d].w[k];
  double s = 0.0;
  for (int p = pop_lowest(x); p >= 0; p = pop_lowest(x)) s = s + (so ? e[p >> 1] : e[p]);
  E0[d] = s;
}

// Lane sums of a 256-thread workgroup in a fixed order.
__device__ double block_sum(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Order 2 from row 0 of H_CI: E[1] = H_00 - E0_0, E[2] = sum_{j>0} H_0j (H_0j / (E0_0 - E0_j)).
__global__ __launch_bounds__(256) void k_mp2_row(const double *__restrict__ row, const double *__restrict__ E0, int ndet,
                                                 double *__restrict__ E) {
  __shared__ double red[256];
  double s = 0.0;
  for (int j = 1 + threadIdx.x; j < ndet; j += 256) s += row[j] * (row[j] / (E0[0] - E0[j]));
  const double t = block_sum(s, red);
  if (threadIdx.x == 0) {
    E[0] = E0[0];
    E[1] = row[0] - E0[0];
    E[2] = t;
  }
}

// t = H c (one wave per row, lanes over the columns, butterfly sum).
__global__ __launch_bounds__(256) void k_matvec(const double *__restrict__ H, size_t ld, int nd, const double *__restrict__ c,
                                                double *__restrict__ t) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= nd) return;
  double s = 0.0;
  for (int j = lane; j < nd; j += 64) s += H[size_t(row) * ld + j] * c[j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) t[row] = s;
}

// The recursion of mp (src/moller_plesset.zig:436-471), step k: E[k] (k = 1: H_00 - E0_0; k > 1:
// sum_{j>0} H_0j C[k-1][j] = t[0], as C[k-1][0] = 0), then for k < order
// C[k][i] = (t[i] - E0_i C[k-1][i] - sum_{j=1..k} E[j] C[k-j][i]) / (E0_0 - E0_i), i > 0, C[k][0] = 0.
__global__ __launch_bounds__(256) void k_mp_step(const double *__restrict__ H, const double *__restrict__ t,
                                                 const double *__restrict__ E0, int nd, int k, int order,
                                                 double *__restrict__ Cc, double *__restrict__ E) {
  // every workgroup forms E[k] itself from the same inputs; only workgroup 0 stores it
  const double ek = k == 1 ? H[0] - E0[0] : t[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) E[k] = ek;
  if (k == order) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nd) return;
  double v = 0.0;
  if (i > 0) {
    const double term2 = E0[i] * Cc[size_t(k - 1) * nd + i];
    double term3 = 0.0;
    for (int j = 1; j <= k; ++j) term3 = term3 + (j == k ? ek : E[j]) * Cc[size_t(k - j) * nd + i];
    v = (t[i] - term2 - term3) / (E0[0] - E0[i]);
  }
  Cc[size_t(k) * nd + i] = v;
}

__global__ void k_unit(double *c, int nd) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nd) c[i] = i == 0 ? 1.0 : 0.0;
}

// generateDets (src/configuration_interaction.zig:126-183): the reference determinant (spin orbitals
// 0..nel-1), then for each requested excitation level k (skipped when 0, above nel or above the
// virtual count) every lexicographic k-subset of the occupied orbitals replaced by every lexicographic
// k-subset of the virtual ones.
bool combos_next(std::vector<int> &c, int n) {
  const int k = int(c.size());
  int i = k - 1;
  while (i >= 0 && c[i] == n - k + i) --i;
  if (i < 0) return false;
  ++c[i];
  for (int j = i + 1; j < k; ++j) c[j] = c[j - 1] + 1;
  return true;
}

std::vector<Det> gen_dets(int nel, int nsp, const int *exc, size_t nexc, size_t limit, bool &too_many) {
  std::vector<Det> out;
  too_many = false;
  Det ref{};
  for (int p = 0; p < nel; ++p) ref.w[p >> 6] |= 1ull << (p & 63);
  out.push_back(ref);
  const int nv = nsp - nel;
  for (size_t x = 0; x < nexc; ++x) {
    const int k = exc[x];
    if (k <= 0 || k > nel || k > nv) continue;
    std::vector<int> oc(k);
    for (int i = 0; i < k; ++i) oc[i] = i;
    do {
      std::vector<int> vc(k);
      for (int i = 0; i < k; ++i) vc[i] = i;
      do {
        if (out.size() >= limit) {
          too_many = true;
          return out;
        }
        Det d = ref;
        for (int i = 0; i < k; ++i) d.w[oc[i] >> 6] &= ~(1ull << (oc[i] & 63));
        for (int i = 0; i < k; ++i) {
          const int p = nel + vc[i];
          d.w[p >> 6] |= 1ull << (p & 63);
        }
        out.push_back(d);
      } while (combos_next(vc, nv));
    } while (combos_next(oc, nel));
  }
  return out;
}

// The spin-orbital setting of one call: the FULL transform of the resident ERIs with C on the device.
struct SoSetup {
  SoInts I{};
  int nel = 0;
};

int so_setup(zq_ctx *ctx, const double *C, int nocc, int generalized, SoSetup &s) {
  if (!ctx->eri) return zq_fail(ctx, ZQ_ERR_STATE, "no ERIs loaded");
  const int n = ctx->n;
  const int nsp = generalized ? n : 2 * n;
  if (nsp > 64 * kWords) return zq_fail(ctx, ZQ_ERR_INVALID, "more than 128 spin orbitals");
  const int nel = generalized ? nocc : 2 * nocc;
  if (nocc < 0 || nel > nsp) return zq_fail(ctx, ZQ_ERR_INVALID, "bad occupation");
  void *d = nullptr;
  int dims[4];
  bool f32 = false;
  if (int rc = zq_transform_device(ctx, C, 0, ZQ_BLOCK_FULL, ZQ_PREC_FP64, &d, dims, &f32)) return rc;
  s.I.g = (const double *)d;
  s.I.n = n, s.I.nsp = nsp, s.I.so = generalized ? 0 : 1;
  s.nel = nel;
  return ZQ_OK;
}

int upload_dets(zq_ctx *ctx, const std::vector<Det> &dets, Det **d) {
  *d = (Det *)zq_scratch(ctx, "ci_dets", dets.size() * sizeof(Det), nullptr, false);
  if (!*d) return zq_fail(ctx, ZQ_ERR_OOM, "determinant list");
  ZQ_TRY(ctx, hipMemcpyAsync(*d, dets.data(), dets.size() * sizeof(Det), hipMemcpyHostToDevice, ctx->stream));
  return ZQ_OK;
}

int build_hci(zq_ctx *ctx, const Det *dd, int nd, int rows, const SoInts &I, double **H) {
  *H = (double *)zq_scratch(ctx, "ci_H", size_t(rows) * nd * sizeof(double), nullptr, false);
  if (!*H) return zq_fail(ctx, ZQ_ERR_OOM, "H_CI");
  zq_prof_begin(ctx, "ci_slater");
  hipLaunchKernelGGL(k_hci, dim3(unsigned((nd + 15) / 16), unsigned((rows + 15) / 16)), dim3(256), 0, ctx->stream, dd, nd,
                     rows, I, *H, size_t(nd));
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, "k_hci");
  return ZQ_OK;
}

}  // namespace

extern "C" int zq_ao2mo_pp(zq_ctx *ctx, size_t n_, const double *A, const double *C, double *out) {
  ZqGate _gate;
  if (!ctx || !A || !C || !out || n_ == 0 || n_ > 8192) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_ao2mo_pp: bad arguments");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  const int n = int(n_);
  const size_t nn = n_ * n_;
  auto *dA = (double *)zq_scratch(ctx, "pp_A", nn * 8, nullptr, false);
  auto *dC = (double *)zq_scratch(ctx, "pp_C", nn * 8, nullptr, false);
  auto *dT = (double *)zq_scratch(ctx, "pp_T", nn * 8, nullptr, false);
  auto *dO = (double *)zq_scratch(ctx, "pp_O", nn * 8, nullptr, false);
  if (!dA || !dC || !dT || !dO) return zq_fail(ctx, ZQ_ERR_OOM, "zq_ao2mo_pp buffers");
  ZQ_TRY(ctx, hipMemcpyAsync(dA, A, nn * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(dC, C, nn * 8, hipMemcpyHostToDevice, ctx->stream));
  // T[mu][q] = sum_nu A[mu][nu] C[nu][q]; out[p][q] = sum_mu C[mu][p] T[mu][q]
  if (int rc = zq_dgemm(ctx, dA, n, 1, dC, n, 1, dT, n, 1, n, n, n, "ao2mo_pp")) return rc;
  if (int rc = zq_dgemm(ctx, dC, 1, n, dT, n, 1, dO, n, 1, n, n, n, "ao2mo_pp")) return rc;
  ZQ_TRY(ctx, hipMemcpyAsync(out, dO, nn * 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ZQ_OK;
}

extern "C" int zq_ci(zq_ctx *ctx, const double *H, const double *C, size_t nocc, int generalized, const int *excitations,
                     size_t nexc, double VN, size_t nstates, double *energies, size_t *ndet_out) {
  ZqGate _gate;
  if (!ctx || !H || !C || !energies || (nexc && !excitations) || nstates == 0)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_ci: bad arguments");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  SoSetup s;
  if (int rc = so_setup(ctx, C, int(nocc), generalized, s)) return rc;
  const int n = s.I.n, nsp = s.I.nsp;
  bool too_many = false;
  const std::vector<Det> dets = gen_dets(s.nel, nsp, excitations, nexc, kMaxDets + 1, too_many);
  if (too_many || dets.size() > kMaxDets)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_ci: more than 16384 determinants (dense H_CI)");
  const int nd = int(dets.size());
  if (ndet_out) *ndet_out = dets.size();
  // h_MO = C^T H C (ao2mo_pp), then in spin orbitals
  const size_t nn = size_t(n) * n;
  auto *dH = (double *)zq_scratch(ctx, "ci_h", nn * 8, nullptr, false);
  auto *dC = (double *)zq_scratch(ctx, "ci_C", nn * 8, nullptr, false);
  auto *dT = (double *)zq_scratch(ctx, "ci_T", nn * 8, nullptr, false);
  auto *dh = (double *)zq_scratch(ctx, "ci_hmo", nn * 8, nullptr, false);
  auto *dhso = (double *)zq_scratch(ctx, "ci_hso", size_t(nsp) * nsp * 8, nullptr, false);
  if (!dH || !dC || !dT || !dh || !dhso) return zq_fail(ctx, ZQ_ERR_OOM, "zq_ci buffers");
  ZQ_TRY(ctx, hipMemcpyAsync(dH, H, nn * 8, hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(dC, C, nn * 8, hipMemcpyHostToDevice, ctx->stream));
  if (int rc = zq_dgemm(ctx, dH, n, 1, dC, n, 1, dT, n, 1, n, n, n, "ao2mo_pp")) return rc;
  if (int rc = zq_dgemm(ctx, dC, 1, n, dT, n, 1, dh, n, 1, n, n, n, "ao2mo_pp")) return rc;
  hipLaunchKernelGGL(k_h_so, dim3(unsigned(nsp)), dim3(128), 0, ctx->stream, dh, n, nsp, s.I.so, dhso);
  ZQ_LAUNCH_CHECK(ctx, "k_h_so");
  s.I.h = dhso;
  Det *dd = nullptr;
  if (int rc = upload_dets(ctx, dets, &dd)) return rc;
  double *Hci = nullptr;
  if (int rc = build_hci(ctx, dd, nd, nd, s.I, &Hci)) return rc;
  auto *w = (double *)zq_scratch(ctx, "ci_w", 2 * size_t(nd) * 8, nullptr, false);
  auto *info = (int *)zq_scratch(ctx, "ci_info", sizeof(int), nullptr, false);
  if (!w || !info) return zq_fail(ctx, ZQ_ERR_OOM, "zq_ci eigen buffers");
  if (int rc = zq_syev_rows(ctx, Hci, nd, w, w + nd, info)) return rc;
  std::vector<double> hw(nd);
  int hinfo = 0;
  ZQ_TRY(ctx, hipMemcpyAsync(hw.data(), w, size_t(nd) * 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (hinfo != 0) return zq_fail(ctx, ZQ_ERR_LAPACK, "zq_ci: eigensolver did not converge");
  for (size_t k = 0; k < nstates && k < size_t(nd); ++k) energies[k] = hw[k] + VN;
  return ZQ_OK;
}

extern "C" int zq_mp(zq_ctx *ctx, const double *C, const double *e, size_t nocc, int generalized, int order,
                     double *energies, size_t *ndet_out) {
  ZqGate _gate;
  if (!ctx || !C || !e || !energies || order < 1 || order > 16) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_mp: bad arguments");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  SoSetup s;
  if (int rc = so_setup(ctx, C, int(nocc), generalized, s)) return rc;
  const int n = s.I.n, nsp = s.I.nsp, nel = s.nel;
  int exc[16], nexc = std::min(order, std::min(nel, nsp - nel));
  for (int i = 0; i < nexc; ++i) exc[i] = i + 1;
  const size_t limit = order == 2 ? kMaxRowDets : kMaxDets;
  bool too_many = false;
  const std::vector<Det> dets = gen_dets(nel, nsp, exc, size_t(nexc), limit + 1, too_many);
  if (too_many || dets.size() > limit)
    return zq_fail(ctx, ZQ_ERR_INVALID, order == 2 ? "zq_mp: more than 2^24 determinants"
                                                   : "zq_mp: more than 16384 determinants (dense H_CI)");
  const int nd = int(dets.size());
  if (ndet_out) *ndet_out = dets.size();
  auto *de = (double *)zq_scratch(ctx, "mp_e", size_t(n) * 8, nullptr, false);
  auto *hms = (double *)zq_scratch(ctx, "mp_hms", size_t(nsp) * nsp * 8, nullptr, false);
  auto *E0 = (double *)zq_scratch(ctx, "mp_E0", size_t(nd) * 8, nullptr, false);
  auto *E = (double *)zq_scratch(ctx, "mp_E", size_t(order + 1) * 8, nullptr, false);
  if (!de || !hms || !E0 || !E) return zq_fail(ctx, ZQ_ERR_OOM, "zq_mp buffers");
  ZQ_TRY(ctx, hipMemcpyAsync(de, e, size_t(n) * 8, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_mp_hms, dim3(unsigned(nsp)), dim3(128), 0, ctx->stream, s.I, de, nel, hms);
  s.I.h = hms;
  Det *dd = nullptr;
  if (int rc = upload_dets(ctx, dets, &dd)) return rc;
  hipLaunchKernelGGL(k_e0, dim3(unsigned((nd + 255) / 256)), dim3(256), 0, ctx->stream, dd, nd, de, s.I.so, E0);
  ZQ_LAUNCH_CHECK(ctx, "k_e0");
  std::vector<double> hE(order + 1, 0.0);
  if (order == 2) {
    double *row = nullptr;
    if (int rc = build_hci(ctx, dd, nd, 1, s.I, &row)) return rc;
    hipLaunchKernelGGL(k_mp2_row, dim3(1), dim3(256), 0, ctx->stream, row, E0, nd, E);
    ZQ_LAUNCH_CHECK(ctx, "k_mp2_row");
  } else {
    double *Hci = nullptr;
    if (int rc = build_hci(ctx, dd, nd, nd, s.I, &Hci)) return rc;
    auto *Cc = (double *)zq_scratch(ctx, "mp_C", size_t(order + 1) * nd * 8, nullptr, false);
    auto *t = (double *)zq_scratch(ctx, "mp_t", size_t(nd) * 8, nullptr, false);
    if (!Cc || !t) return zq_fail(ctx, ZQ_ERR_OOM, "zq_mp buffers");
    const unsigned g256 = unsigned((nd + 255) / 256);
    hipLaunchKernelGGL(k_unit, dim3(g256), dim3(256), 0, ctx->stream, Cc, nd);
    for (int k = 1; k <= order; ++k) {
      hipLaunchKernelGGL(k_matvec, dim3(unsigned((nd + 3) / 4)), dim3(256), 0, ctx->stream, Hci, size_t(nd), nd,
                         Cc + size_t(k - 1) * nd, t);
      hipLaunchKernelGGL(k_mp_step, dim3(g256), dim3(256), 0, ctx->stream, Hci, t, E0, nd, k, order, Cc, E);
    }
    ZQ_LAUNCH_CHECK(ctx, "k_mp_step");
  }
  ZQ_TRY(ctx, hipMemcpyAsync(hE.data() + 1, E + 1, size_t(order) * 8, hipMemcpyDeviceToHost, ctx->stream));
  double e00 = 0.0;
  ZQ_TRY(ctx, hipMemcpyAsync(&e00, E0, 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  hE[0] = e00;
  for (int k = 0; k <= order; ++k) energies[k] = hE[k];
  return ZQ_OK;
}
