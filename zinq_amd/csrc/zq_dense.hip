// Small dense FP64 kernels of the device SCF: an n x n GEMM sized for latency and the SP2 density
// purification that replaces the per-iteration diagonalisation.
//
// zq_gemm_nt: C = alpha A B^T (A: M x K, B: N x K, both read along rows).  One 64-lane wave owns a
//   16 x 16 tile of C and walks K in chunks of 16 with v_mfma_f64_16x16x4f64.  Inside a chunk lane l
//   (r = l % 16, q = l / 16) holds A[i0+r][k0+4q .. k0+4q+3] and B[j0+r][same k]: the 4 MFMAs of the
//   chunk take element t of every lane, i.e. the k order inside a chunk is permuted, which leaves the
//   sum unchanged and turns every operand fetch into one 32-byte row segment.  The next 64 k are
//   loaded before the current 64 are multiplied.  No LDS, no barriers: the kernel is latency-bound at
//   the SCF's sizes (n <= a few hundred), so the grid has one small workgroup per tile.
//   Every operand of the SCF is symmetric or stored transposed, so A B^T covers all of its products.
//
// SP2 (second-order spectral projection, A. M. N. Niklasson, Phys. Rev. B 66, 155115 (2002)):
//   X_0 = (e_max I - F') / (e_max - e_min)    Gershgorin bounds of the orthonormal-basis Fock matrix
//   X_{k+1} = X_k^2        if |tr(X_k^2) - N| <= |2 tr(X_k) - tr(X_k^2) - N|
//             2 X_k - X_k^2 otherwise
//   converges to the projector onto the N lowest eigenvectors of F' (the aufbau density) when F' has
//   a gap between eigenvalues N and N+1.  idem_k = tr(X_k) - tr(X_k^2) = sum lambda (1 - lambda) >= 0
//   bounds every eigenvalue's distance to {0, 1}; once idem_k < tol two more steps square every
//   eigenvalue's distance at least once, and the iteration stops.
//   X symmetric, so only tiles ti <= tj are computed and mirrored; the epilogue applies the branch and
//   writes per-tile partial traces (tr X, ||X||_F^2 = tr X^2), which every workgroup of the next step
//   sums in the same fixed order, so all of them take the same branch.  X is ping-ponged between two
//   buffers (input [p], output [p^1]).
//   Two forms, bitwise equal on plain runs: k_sp2_run takes every step in one launch (tiles handed out
//   by ticket, a step's tiles counted before the next reads X), which the device SCF uses; k_sp2_step
//   is one launch per step, used by the host-checked driver.
//
// Certified scale-and-fold steps (one-launch form in the device SCF).  With bounds a_k >= every
// virtual eigenvalue of X_k and b_k <= every occupied one (a_k < b_k), the down step may fold at
// s = a_k / 2:  X <- ((X - s I) / (1 - s))^2, which maps [0, a_k] into [0, f(a_k)] and keeps [b_k, 1]
// above it, and the up step at s = (1 - b_k) / 2:  X <- 2 X / (1 - s) - X^2 / (1 - s)^2, its mirror
// image.  Each step folds the side whose bound lags (down when a_k > 1 - b_k); the gap widens about
// twice as fast per step as with X^2 / 2X - X^2, both bounds follow the same maps, and as they approach
// {0, 1} the fold points vanish.  The bounds come from an earlier
// plain run of the same SCF: its traces at the first step with tr X - tr X^2 < 1e-3 give
// sum_virt x <= ((idem + tr X - N) / 2 + 4 idem^2) and sum_occ (1 - x) <= ((idem - tr X + N) / 2 + 4 idem^2),
// hence the LUMO's image from above and the HOMO's from below, which the inverted branch maps carry back
// to energies (a lower bound of the LUMO, an upper bound of the HOMO of that run's F').  By Weyl's
// inequality no eigenvalue of the new F' lies further than D = ||F' - F'_ref||_F from the old one, so
// LUMO - D and HOMO + D bound the new run; it folds only when the gap left, (LUMO - D) - (HOMO + D),
// exceeds half the estimated gap, and runs plain otherwise (becoming the new reference).

#include <algorithm>

#include "zq_internal.h"
#include "zq_tile.h"

namespace {

using namespace zqt;

// Sp2Bufs::est, in doubles: the fold bookkeeping kept across the SP2 runs of one device SCF.
enum : int {
  kEstValid = 0,      // 1: ref holds F' of a settled plain run, and kEstHomo / kEstLumo bound its HOMO / LUMO
  kEstHomo = 1,       // upper bound of that run's HOMO energy
  kEstLumo = 2,       // lower bound of its LUMO energy
  kEstLastPlain = 3,  // 1: the last run was plain: ref holds its F', its traces are in the history
  kEstLastLo = 4,     // and its Gershgorin interval
  kEstLastHi = 5,
  kEstFold = 6,       // this run folds, with the bounds kEstA0 (virtual) and kEstB0 (occupied) in X_0's scale
  kEstA0 = 7,
  kEstB0 = 8,
  kEstSteps = 9,      // steps the last run recorded, and whether it settled
  kEstSettled = 10,
  kEstRuns = 11,      // since the SCF's first density: densities, folded ones, SP2 steps in all
  kEstFolds = 12,
  kEstStepSum = 13,
  kEstRefSel = 14,    // which of the two n x n reference buffers holds F'_ref
  kEstHist = 16,      // history[k] = {tr X_k, tr X_k^2, 1 for a down step, fold point}
  kEstSize = kEstHist + 4 * kSp2MaxSteps
};

template <bool VEC>
__global__ __launch_bounds__(64) void k_gemm_nt(const double *__restrict__ A, int lda, const double *__restrict__ B,
                                                int ldb, double *__restrict__ C, int ldc, int M, int N, int K,
                                                double alpha, const int *__restrict__ skip, const int *__restrict__ bsel,
                                                long long bstride) {
  if (skip && *skip) return;
  if (bsel) B += *bsel * bstride;
  const int i0 = blockIdx.y * 16, j0 = blockIdx.x * 16;
  const d4 acc = tile_nt<VEC>(A, lda, B, ldb, i0, j0, M, N, K);
  const int lane = threadIdx.x, col = j0 + (lane & 15);
  if (col >= N) return;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int row = i0 + 4 * j + (lane >> 4);
    if (row < M) C[size_t(row) * ldc + col] = alpha * acc[j];
  }
}

__device__ inline void tile_of(int t, int T, int &ti, int &tj) {
  // t enumerates the pairs ti <= tj row by row: row ti holds T - ti tiles
  ti = 0;
  while (t >= T - ti) {
    t -= T - ti;
    ++ti;
  }
  tj = ti + t;
}

// HOMO (upper bound) and LUMO (lower bound) energies of the last plain run (steps steps, Gershgorin
// interval [lo, hi]) from its trace history h (a copy in LDS of est + kEstHist).
__device__ bool sp2_estimate(int steps, double lo, double hi, const double *h, int nocc, double &eh, double &el) {
  int ks = -1;
  for (int k = 0; k < steps && ks < 0; ++k)
    if (h[4 * k] - h[4 * k + 1] < 1e-3) ks = k;
  if (ks < 0) return false;
  const double idem = h[4 * ks] - h[4 * ks + 1], d = h[4 * ks] - double(nocc), r = 4.0 * idem * idem + 1e-12;
  double xl = fmin(fmax(0.5 * (idem + d) + r, 0.0), 1.0), xh = fmin(fmax(1.0 - (0.5 * (idem - d) + r), 0.0), 1.0);
  for (int k = ks - 1; k >= 0; --k) {
    if (h[4 * k + 2] != 0.0) {
      xl = sqrt(xl), xh = sqrt(xh);
    } else {
      xl = 1.0 - sqrt(fmax(1.0 - xl, 0.0)), xh = 1.0 - sqrt(fmax(1.0 - xh, 0.0));
    }
  }
  const double inv = 1.0 / fmax(hi - lo, 1e-10 * fmax(1.0, fabs(hi)));
  eh = hi - xh / inv;
  el = hi - xl / inv;
  return el > eh;
}

// Gershgorin bounds of the symmetric n x n F (ld n): scal[0] = min_i (F_ii - R_i), scal[1] = max_i (F_ii + R_i).
// kGershBlocks workgroups of 4 waves, one wave per row (global wave g takes rows g, g + 4 kGershBlocks,
// ...): the lanes read the row coalesced and sum it with a butterfly.  Each workgroup leaves its
// partials in gpart and counts itself in *gcount after a release fence; the last one, after an acquire
// fence, combines them in workgroup order and clears the count.  track != 0 (one-launch SP2 of the
// device SCF; 1 starts an SCF) also decides whether this run folds: the last workgroup turns the last
// plain run's history into bounds, bounds the shift ||F - ref||_2 by the smaller of ||F - ref||_F and
// the largest row sum of |F - ref| (ref: the buffer kEstRefSel names), and writes kEstFold / kEstA0 /
// kEstB0.  The same pass stores F into the other buffer; a plain run
// makes that one the reference.
constexpr int kGershBlocks = 16;

__global__ __launch_bounds__(256) void k_gershgorin(const double *__restrict__ F, int n, double *__restrict__ scal,
                                                    const int *__restrict__ skip, int track, int nocc,
                                                    double *__restrict__ est, double *__restrict__ ref,
                                                    double *__restrict__ gpart, unsigned *__restrict__ gcount) {
  __shared__ double lo_s[4], hi_s[4], dd_s[4], dm_s[4], hist_s[4 * kSp2MaxSteps];
  __shared__ int last_s;
  if (skip && *skip) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // the bookkeeping scalars are read before the rows, so that their latency hides behind the sums
  double e_valid = 0.0, e_homo = 0.0, e_lumo = 0.0, e_lo = 0.0, e_hi = 0.0, e_runs = 0.0, e_folds = 0.0;
  int e_steps = 0, sel = 0;
  bool last_plain = false, pending = false;
  if (track) {
    e_valid = est[kEstValid], e_homo = est[kEstHomo], e_lumo = est[kEstLumo];
    e_lo = est[kEstLastLo], e_hi = est[kEstLastHi], e_steps = int(est[kEstSteps]);
    sel = int(est[kEstRefSel]) & 1;
    e_runs = est[kEstRuns], e_folds = est[kEstFolds];
    last_plain = est[kEstLastPlain] != 0.0;
    pending = track == 2 && last_plain && est[kEstSettled] != 0.0;
  }
  const double *rold = ref + size_t(sel) * n * n;
  double *rnew = ref + size_t(sel ^ 1) * n * n;
  double lo = 1e300, hi = -1e300, dd = 0.0, dm = 0.0;  // dm: the largest row sum of |F' - F'_ref|
  const int gw = blockIdx.x * 4 + w, ngw = gridDim.x * 4;
  for (int i = gw; i < n; i += ngw) {
    const double *row = F + size_t(i) * n;
    const double d = row[i];
    double rsum = 0.0, rdif = 0.0;
    for (int j0 = 0; j0 < n; j0 += 256) {
      double f[4], r[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = j0 + 64 * u + lane;
        f[u] = j < n ? row[j] : 0.0;
        r[u] = (track == 2 && j < n) ? rold[size_t(i) * n + j] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = j0 + 64 * u + lane;
        rsum += fabs(f[u]);
        const double df = f[u] - r[u];
        if (track == 2) dd += df * df, rdif += fabs(df);
        if (track && j < n) rnew[size_t(i) * n + j] = f[u];
      }
    }
    rsum = wave_sum(rsum) - fabs(d);
    lo = fmin(lo, d - rsum);
    hi = fmax(hi, d + rsum);
    if (track == 2) dm = fmax(dm, wave_sum(rdif));
  }
  dd = wave_sum(dd);
  if (lane == 0) lo_s[w] = lo, hi_s[w] = hi, dd_s[w] = dd, dm_s[w] = dm;
  __syncthreads();
  if (threadIdx.x == 0) {
    gpart[4 * blockIdx.x] = fmin(fmin(lo_s[0], lo_s[1]), fmin(lo_s[2], lo_s[3]));
    gpart[4 * blockIdx.x + 1] = fmax(fmax(hi_s[0], hi_s[1]), fmax(hi_s[2], hi_s[3]));
    gpart[4 * blockIdx.x + 2] = ((dd_s[0] + dd_s[1]) + dd_s[2]) + dd_s[3];
    gpart[4 * blockIdx.x + 3] = fmax(fmax(dm_s[0], dm_s[1]), fmax(dm_s[2], dm_s[3]));
    __threadfence();
    last_s = atomicAdd(gcount, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last_s) return;
  __threadfence();
  // the last plain run's history into LDS for the estimate
  if (pending)
    for (int e = threadIdx.x; e < 4 * e_steps; e += blockDim.x) hist_s[e] = est[kEstHist + e];
  __syncthreads();
  if (threadIdx.x == 0) {
    *gcount = 0u;
    double glo = gpart[0], ghi = gpart[1], dsum = gpart[2], dmax = gpart[3];
    for (int k = 1; k < int(gridDim.x); ++k)
      glo = fmin(glo, gpart[4 * k]), ghi = fmax(ghi, gpart[4 * k + 1]), dsum += gpart[4 * k + 2],
      dmax = fmax(dmax, gpart[4 * k + 3]);
    // widen by a relative margin: the row sums above carry rounding errors
    const double pad = 1e-12 * fmax(fabs(glo), fabs(ghi));
    const double emin = glo - pad, emax = ghi + pad;
    scal[0] = emin;
    scal[1] = emax;
    int fold = 0;
    if (track) {
      double valid = track == 1 ? 0.0 : e_valid, eh = e_homo, el = e_lumo;
      if (track == 2 && last_plain) {
        // ref holds the last run's F': its history gives the bounds, when it settled
        valid = 0.0;
        if (pending && sp2_estimate(e_steps, e_lo, e_hi, hist_s, nocc, eh, el)) {
          est[kEstHomo] = eh, est[kEstLumo] = el;
          valid = 1.0;
        }
      }
      est[kEstValid] = valid;
      if (valid != 0.0) {
        // ||F' - F'_ref||_2 is at most its Frobenius norm and at most its largest absolute row sum
        // (F' symmetric); the row sum is often the smaller one early in an SCF.  A relative pad covers
        // the rounding of both sums.
        const double D = fmin(sqrt(dsum), dmax) * (1.0 + 1e-12);
        const double lumo = el - D, homo = eh + D;
        if (lumo - homo > 0.5 * (el - eh)) {
          const double inv = 1.0 / fmax(emax - emin, 1e-10 * fmax(1.0, fabs(emax)));
          fold = 1;
          est[kEstA0] = fmin(fmax((emax - lumo) * inv, 0.0), 1.0);
          est[kEstB0] = fmin(fmax((emax - homo) * inv, 0.0), 1.0);
        }
      }
      est[kEstFold] = fold;
      est[kEstLastPlain] = fold ? 0.0 : 1.0;
      if (!fold) est[kEstRefSel] = double(sel ^ 1);
      if (track == 1) e_runs = 0.0, e_folds = 0.0, est[kEstStepSum] = 0.0;
      est[kEstRuns] = e_runs + 1.0;
      est[kEstFolds] = e_folds + fold;
      if (!fold) est[kEstLastLo] = emin, est[kEstLastHi] = emax;
    }
  }
}

// Writes the partial traces of the tile this wave holds (vals[j] at rows i0+4j+q, column j0+r).
__device__ inline void tile_partials(const double v[4], int ti, int tj, int i0, int j0, double2 *__restrict__ part, int t) {
  const int lane = threadIdx.x, r = lane & 15, q = lane >> 4;
  double s1 = 0.0, s2 = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (ti == tj && i0 + 4 * j + q == j0 + r) s1 += v[j];
    s2 += v[j] * v[j];
  }
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if (lane == 0) part[t] = make_double2(s1, ti == tj ? s2 : 2.0 * s2);
}

// X_0 on the m x m padded grid (zero outside n x n), its partials, the stop state, and the zeroed
// counters of the one-launch form.
__global__ __launch_bounds__(64) void k_sp2_init(const double *__restrict__ F, int n, int m, const double *__restrict__ scal,
                                                 double *__restrict__ X, double2 *__restrict__ part, int *__restrict__ state,
                                                 unsigned *__restrict__ sync, const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int T = m / 16;
  int ti, tj;
  tile_of(blockIdx.x, T, ti, tj);
  const int i0 = ti * 16, j0 = tj * 16, lane = threadIdx.x, r = lane & 15, q = lane >> 4;
  const double emin = scal[0], emax = scal[1];
  const double inv = 1.0 / fmax(emax - emin, 1e-10 * fmax(1.0, fabs(emax)));
  double v[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int row = i0 + 4 * j + q, col = j0 + r;
    double x = 0.0;
    if (row < n && col < n) x = ((row == col ? emax : 0.0) - F[size_t(row) * n + col]) * inv;
    v[j] = x;  // the mirror gets the same value: X_0 is exactly symmetric
    X[size_t(row) * m + col] = x;
    if (ti != tj) X[size_t(col) * m + row] = x;
  }
  tile_partials(v, ti, tj, i0, j0, part, blockIdx.x);
  if (blockIdx.x == 0) {
    for (int u = lane; u < 2 * kSp2MaxSteps; u += 64) sync[u] = 0u;
    if (lane == 0) state[0] = -1, state[1] = 0, state[4] = 0, state[5] = 0;
  }
}

// The operands of one SP2 tile (ti <= tj).  For m <= 256 (PRE) the workgroup's four waves split K by
// 16-wide chunks, wave w taking chunks w, w + 4, w + 8, w + 12 below m, and each loads all of its chunks
// at once: the tile costs one memory round trip and a quarter of the MFMA chain, and no MFMA multiplies
// the zero padding above m.  Larger m: one wave walks K with the double-buffered tile_nt.  Wave 0 also
// holds its own X_in elements, for 2 X - X^2.
template <bool PRE>
struct Sp2Ops {
  double x[4];
  Frag f;
};

constexpr int sp2_threads(bool pre) { return pre ? 256 : 64; }

template <bool PRE>
__device__ inline void sp2_load(Sp2Ops<PRE> &o, const double *__restrict__ Xin, int m, int i0, int j0) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4, w = threadIdx.x >> 6;
  if (w == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) o.x[j] = Xin[size_t(i0 + 4 * j + q) * m + j0 + r];
  }
  if (PRE) {
    // m % 16 == 0 and m is even: a chunk lies wholly inside or wholly outside, and rows are 16-byte aligned
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = w + 4 * i;
      if (16 * c < m) {
        const double2 *pa = reinterpret_cast<const double2 *>(Xin + size_t(i0 + r) * m + 16 * c + 4 * q);
        const double2 *pb = reinterpret_cast<const double2 *>(Xin + size_t(j0 + r) * m + 16 * c + 4 * q);
        const double2 a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
        o.f.a[i][0] = a0.x, o.f.a[i][1] = a0.y, o.f.a[i][2] = a1.x, o.f.a[i][3] = a1.y;
        o.f.b[i][0] = b0.x, o.f.b[i][1] = b0.y, o.f.b[i][2] = b1.x, o.f.b[i][3] = b1.y;
      }
    }
  }
}

// X_out tile = alpha X_in^2 + beta X_in + gamma I, stored with its mirror, and its partial traces in
// pout[t].  Every wave of the workgroup calls it; PRE sums the waves' partial products in LDS, in wave
// order, and wave 0 finishes the tile.
// The polynomial of one step, X_(k+1) = alpha X_k^2 + beta X_k + gamma I; s is the fold point (0 plain).
struct Sp2Map {
  double alpha, beta, gamma, s;
  bool down;
};

template <bool PRE>
__device__ inline void sp2_finish(const Sp2Ops<PRE> &o, const double *__restrict__ Xin, double *__restrict__ Xout, int m,
                                  int ti, int tj, const Sp2Map &p, double2 *__restrict__ pout, int t) {
  const int i0 = ti * 16, j0 = tj * 16, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4, w = threadIdx.x >> 6;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  if (PRE) {
    __shared__ d4 red[3][64];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (16 * (w + 4 * i) < m)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(o.f.a[i][u], o.f.b[i][u], acc, 0, 0, 0);
    if (w > 0) red[w - 1][lane] = acc;
    __syncthreads();
    if (w == 0) {
      acc += red[0][lane];
      acc += red[1][lane];
      acc += red[2][lane];
    }
  } else {
    acc = tile_nt<true>(Xin, m, Xin, m, i0, j0, m, m, m);
  }
  if (w != 0) return;
  double v[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    // plain steps: exactly X^2 (alpha 1, beta 0) or 2X - X^2 (alpha -1, beta 2), one rounding
    v[j] = fma(p.alpha, acc[j], p.beta * o.x[j]);
    if (i0 + 4 * j + q == j0 + r) v[j] += p.gamma;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int row = i0 + 4 * j + q, col = j0 + r;
    Xout[size_t(row) * m + col] = v[j];
    if (ti != tj) Xout[size_t(col) * m + row] = v[j];
  }
  tile_partials(v, ti, tj, i0, j0, pout, t);
}

// The stop rule and the step's map from the summed partials (t1 = tr X_k, t2 = tr X_k^2 over the m x m
// grid): count < 0 before idem < tol was first seen, then the number of steps taken since; the step that
// brings count to 2 produces the settled projector.  Plain: X^2 or 2X - X^2, whichever keeps the trace
// closer to N.  Folding (bounds a < b carried in a, b): the folded map of the side whose bound lags,
// down when a > 1 - b, else up; both bounds go through it.  (The trace rule cannot choose between
// folded maps: a folded down step lifts every eigenvalue near 0, the zero padding included, to
// s^2/(1-s)^2, which makes its trace look worse than it is and keeps the iteration going up.)
__device__ inline Sp2Map sp2_map(double t1, double t2, double tol, int nocc, int &count, bool fold, double &a,
                                 double &b) {
  const double N = double(nocc);
  if (count < 0 && t1 - t2 < tol) count = 0;
  if (count >= 0) ++count;
  Sp2Map p;
  if (!fold || !(b > a)) {
    p.down = fabs(t2 - N) <= fabs(2.0 * t1 - t2 - N);
    p.alpha = p.down ? 1.0 : -1.0, p.beta = p.down ? 0.0 : 2.0, p.gamma = 0.0, p.s = 0.0;
    return p;
  }
  const double sd = 0.5 * a, su = 0.5 * (1.0 - b);
  const double cd = 1.0 / (1.0 - sd), cu = 1.0 / (1.0 - su);
  p.down = a > 1.0 - b;
  if (p.down)
    p.alpha = cd * cd, p.beta = -2.0 * sd * cd * cd, p.gamma = sd * sd * cd * cd, p.s = sd;
  else
    p.alpha = -cu * cu, p.beta = 2.0 * cu, p.gamma = 0.0, p.s = su;
  a = p.alpha * a * a + p.beta * a + p.gamma;
  b = p.alpha * b * b + p.beta * b + p.gamma;
  return p;
}

__device__ inline void sum_partials(const double2 *__restrict__ pin, int ntiles, double &t1, double &t2) {
  t1 = 0.0, t2 = 0.0;
  for (int u = threadIdx.x & 63; u < ntiles; u += 64) t1 += pin[u].x, t2 += pin[u].y;
  t1 = wave_sum(t1);
  t2 = wave_sum(t2);
}

// One SP2 step: X_out from X_in (m x m, symmetric), partials and state from [p] to [p^1].
// state = {count, steps}; count >= 2 means settled.  The step that brings count to 2 stores the parity
// of its output buffer in *final; later steps only carry the state forward and return.
template <bool PRE>
__global__ __launch_bounds__(sp2_threads(PRE)) void k_sp2_step(const double *__restrict__ Xin, double *__restrict__ Xout, int m, int nocc,
                                                 const double2 *__restrict__ pin, double2 *__restrict__ pout,
                                                 const int *__restrict__ sin, int *__restrict__ sout, int *__restrict__ final,
                                                 int pout_parity, double tol, const int *__restrict__ skip) {
  if (skip && *skip) return;
  if (sin[0] >= 2) {  // settled: carry the state forward
    if (blockIdx.x == 0 && threadIdx.x == 0) sout[0] = sin[0], sout[1] = sin[1];
    return;
  }
  const int T = m / 16, ntiles = T * (T + 1) / 2;
  int ti, tj;
  tile_of(blockIdx.x, T, ti, tj);
  // issue every load first: state, this tile's operands, then the partials
  int count = sin[0];
  const int steps = sin[1];
  Sp2Ops<PRE> o;
  sp2_load<PRE>(o, Xin, m, ti * 16, tj * 16);
  double t1, t2, a = 0.0, b = 1.0;
  sum_partials(pin, ntiles, t1, t2);
  const Sp2Map mp = sp2_map(t1, t2, tol, nocc, count, false, a, b);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sout[0] = count, sout[1] = steps + 1;
    if (count == 2) *final = pout_parity;  // X_out is the settled projector
  }
  sp2_finish<PRE>(o, Xin, Xout, m, ti, tj, mp, pout, blockIdx.x);
}

// Bounded waits of the one-launch SP2: a wait that runs out raises *abort, and so makes the launch's
// result count as unsettled; a wait that sees *abort raised gives up too.
constexpr unsigned kSpinLimit = 1u << 22;

// Tiles per workgroup and step of the one-launch SP2 (the grid is ceil(tiles / kSp2TilesPerGroup)).
constexpr int kSp2TilesPerGroup = 1;

// With several waves a workgroup (NW > 1) takes its tiles through wave 0: lane 0 of wave 0 draws the
// ticket and hands it over in LDS, wave 0 alone publishes (it alone writes the tile) and waits, and a
// barrier hands its acquire on to the other waves.  The ticket slot alternates: slot s is written again
// two claims later, after the barrier of the claim in between, which every wave passes only after it
// has read slot s.
template <int NW>
__device__ inline unsigned claim(unsigned *ticket, int &slot) {
  if (NW == 1) {
    unsigned t = 0;
    if (threadIdx.x == 0) t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __shfl(t, 0, 64);
  }
  __shared__ unsigned s_ticket[2];
  if (threadIdx.x == 0) s_ticket[slot] = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const unsigned t = s_ticket[slot];
  slot ^= 1;
  return t;
}

// Counts a workgroup's finished tiles of a step (n of them).  Every wave has read X_in before each tile's
// reduction barrier, so only the writer, wave 0, orders its stores before the count.  The release fence
// writes back the XCD's L2 and is the costliest part of a step's hand-over, so it is issued once per
// workgroup and step, not once per tile.
template <int NW>
__device__ inline void publish(unsigned *done, unsigned n) {
  if (NW == 1) __syncthreads();
  if (threadIdx.x < 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (threadIdx.x == 0) __hip_atomic_fetch_add(done, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <int NW>
__device__ inline bool wait_done(const unsigned *done, unsigned target, int *abort) {
  __shared__ int s_ok;
  bool ok = true;
  if (threadIdx.x < 64) {
    unsigned spins = 0;
    while (__hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if ((spins & 63) == 63 && __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        ok = false;
        break;
      }
      if (++spins > kSpinLimit) {
        __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = false;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (NW > 1 && threadIdx.x == 0) s_ok = ok;
  }
  __syncthreads();
  if (NW > 1) ok = s_ok;
  return ok;
}

// Every step of SP2 in one launch.  The tiles of step k are handed out by the ticket counter sync[k];
// a workgroup takes tiles until none is left, counts each finished one in sync[kSp2MaxSteps + k], and
// waits for all of step k before it reads X_(k+1).  A wait only ever depends on tiles that running
// workgroups have taken, so the launch finishes whichever of its workgroups are resident.  Partials are
// kept per step (part[k] holds those of X_k), so a workgroup that starts late replays the branch and
// stop decisions of the steps it missed from them; every workgroup sums them in the same order and so
// takes the same decisions.  Each tile is computed as by k_sp2_step, so the result is bitwise that of
// the step-per-launch form.  state gets {count, steps} in the pair 2 * (nsteps & 1) and the settled
// buffer's parity in [4], as k_sp2_step leaves them; [5] is the abort flag.
template <bool PRE>
__global__ __launch_bounds__(sp2_threads(PRE)) void k_sp2_run(double *__restrict__ X, size_t mm, int m, int nocc,
                                                double2 *__restrict__ part, unsigned *__restrict__ sync,
                                                int *__restrict__ state, int nsteps, double tol,
                                                const int *__restrict__ skip, double *__restrict__ est) {
  if (skip && *skip) return;
  const int T = m / 16, ntiles = T * (T + 1) / 2;
  unsigned *ticket = sync, *done = sync + kSp2MaxSteps;
  // folding bounds, carried identically by every wave; the history is kept by block 0
  bool fold = false;
  double fa = 0.0, fb = 1.0;
  if (est && est[kEstFold] != 0.0) fold = true, fa = est[kEstA0], fb = est[kEstB0];
  double *hist = (est && blockIdx.x == 0 && threadIdx.x == 0) ? est + kEstHist : nullptr;
  int *abort = state + 5;
  constexpr int NW = sp2_threads(PRE) / 64;
  int count = -1, steps = 0, fin = 0, slot = 0;
  unsigned t = claim<NW>(&ticket[0], slot);
  for (int k = 0; k < nsteps; ++k) {
    const double *Xin = X + (k & 1) * mm;
    double *Xout = X + ((k + 1) & 1) * mm;
    // the operands of the tile already taken are loaded before the partials, so that both loads overlap
    Sp2Ops<PRE> o;
    int ti = 0, tj = 0;
    if (t < unsigned(ntiles)) {
      tile_of(int(t), T, ti, tj);
      sp2_load<PRE>(o, Xin, m, ti * 16, tj * 16);
    }
    double t1, t2;
    sum_partials(part + size_t(k) * ntiles, ntiles, t1, t2);
    const Sp2Map mp = sp2_map(t1, t2, tol, nocc, count, fold, fa, fb);
    if (hist) hist[4 * k] = t1, hist[4 * k + 1] = t2, hist[4 * k + 2] = mp.down ? 1.0 : 0.0, hist[4 * k + 3] = mp.s;
    steps = k + 1;
    if (count == 2) fin = (k + 1) & 1;
    double2 *pout = part + size_t(k + 1) * ntiles;
    unsigned mine = 0;
    while (t < unsigned(ntiles)) {
      sp2_finish<PRE>(o, Xin, Xout, m, ti, tj, mp, pout, int(t));
      ++mine;
      t = claim<NW>(&ticket[k], slot);
      if (t < unsigned(ntiles)) {
        tile_of(int(t), T, ti, tj);
        sp2_load<PRE>(o, Xin, m, ti * 16, tj * 16);
      }
    }
    if (mine) publish<NW>(&done[k], mine);
    if (count >= 2 || k + 1 == nsteps) break;  // the launch's end publishes the last step
    t = claim<NW>(&ticket[k + 1], slot);
    if (!wait_done<NW>(&done[k], unsigned(ntiles), abort)) break;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    int *pair = state + 2 * (nsteps & 1);
    pair[0] = count, pair[1] = steps;
    state[4] = fin;
    if (est) est[kEstSteps] = steps, est[kEstSettled] = count >= 2 ? 1.0 : 0.0, est[kEstStepSum] += steps;
  }
}

}  // namespace

int zq_gemm_nt(zq_ctx *ctx, const double *A, int lda, const double *B, int ldb, double *C, int ldc, int M, int N, int K,
               double alpha, const char *name, const int *skip, const int *bsel, long long bstride) {
  if (M <= 0 || N <= 0 || K <= 0) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_gemm_nt: empty product");
  if (lda < K || ldb < K || ldc < N) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_gemm_nt: leading dimension too small");
  const bool vec = K % 16 == 0 && lda % 2 == 0 && ldb % 2 == 0 && bstride % 2 == 0 &&
                   (reinterpret_cast<uintptr_t>(A) & 15) == 0 && (reinterpret_cast<uintptr_t>(B) & 15) == 0;
  const dim3 grid((N + 15) / 16, (M + 15) / 16);
  zq_prof_begin(ctx, name);
  if (vec)
    hipLaunchKernelGGL(k_gemm_nt<true>, grid, dim3(64), 0, ctx->stream, A, lda, B, ldb, C, ldc, M, N, K, alpha, skip, bsel,
                       bstride);
  else
    hipLaunchKernelGGL(k_gemm_nt<false>, grid, dim3(64), 0, ctx->stream, A, lda, B, ldb, C, ldc, M, N, K, alpha, skip, bsel,
                       bstride);
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, "k_gemm_nt");
  return ZQ_OK;
}

namespace {

// X: two m x m buffers; part: the partials of X_0 .. X_kSp2MaxSteps (ntiles each; the step-per-launch
// form ping-pongs between the first two); sync: the tickets and finished counts of the one-launch form.
struct Sp2Bufs {
  int m, T, ntiles;
  size_t mm;
  double *X, *scal, *est, *ref, *gpart;
  double2 *part;
  unsigned *sync, *gcount;
};

int sp2_bufs(zq_ctx *ctx, int n, Sp2Bufs &b) {
  b.m = (n + 15) / 16 * 16, b.T = b.m / 16, b.ntiles = b.T * (b.T + 1) / 2;
  b.mm = size_t(b.m) * b.m;
  b.X = (double *)zq_scratch(ctx, "sp2_X", 2 * b.mm * sizeof(double));
  b.part = (double2 *)zq_scratch(ctx, "sp2_part", size_t(kSp2MaxSteps + 1) * b.ntiles * sizeof(double2));
  b.scal = (double *)zq_scratch(ctx, "sp2_scal", 2 * sizeof(double));
  b.sync = (unsigned *)zq_scratch(ctx, "sp2_sync", 2 * kSp2MaxSteps * sizeof(unsigned));
  b.est = (double *)zq_scratch(ctx, "sp2_est", kEstSize * sizeof(double));
  b.ref = (double *)zq_scratch(ctx, "sp2_ref", 2 * size_t(n) * n * sizeof(double));
  b.gpart = (double *)zq_scratch(ctx, "sp2_gpart", 4 * kGershBlocks * sizeof(double));
  bool fresh = false;  // the Gershgorin arrival count starts at zero and its last workgroup clears it again
  b.gcount = (unsigned *)zq_scratch(ctx, "sp2_gcount", sizeof(unsigned), &fresh);
  if (!b.X || !b.part || !b.scal || !b.sync || !b.est || !b.ref || !b.gpart || !b.gcount)
    return zq_fail(ctx, ZQ_ERR_OOM, "SP2 buffers");
  if (fresh) ZQ_TRY(ctx, hipMemsetAsync(b.gcount, 0, sizeof(unsigned), ctx->stream));
  return ZQ_OK;
}

// track: 0 plain SP2 outside the fold bookkeeping, 1 the first run of a device SCF, 2 a later run.
void sp2_start(zq_ctx *ctx, const Sp2Bufs &b, const double *dF, int n, int nocc, int *state, const int *skip,
               int track) {
  hipLaunchKernelGGL(k_gershgorin, dim3(kGershBlocks), dim3(256), 0, ctx->stream, dF, n, b.scal, skip, track, nocc, b.est,
                     b.ref, b.gpart, b.gcount);
  hipLaunchKernelGGL(k_sp2_init, dim3(b.ntiles), dim3(64), 0, ctx->stream, dF, n, b.m, b.scal, b.X, b.part, state, b.sync,
                     skip);
}

void sp2_step(zq_ctx *ctx, const Sp2Bufs &b, int k, int nocc, double tol, int *state, const int *skip) {
  const int p = k & 1;
  double *Xi = b.X + p * b.mm, *Xo = b.X + (p ^ 1) * b.mm;
  const double2 *pi = b.part + p * b.ntiles;
  double2 *po = b.part + (p ^ 1) * b.ntiles;
  if (b.m <= 256)
    hipLaunchKernelGGL(k_sp2_step<true>, dim3(b.ntiles), dim3(sp2_threads(true)), 0, ctx->stream, Xi, Xo, b.m, nocc, pi, po,
                       state + 2 * p, state + 2 * (p ^ 1), state + 4, p ^ 1, tol, skip);
  else
    hipLaunchKernelGGL(k_sp2_step<false>, dim3(b.ntiles), dim3(64), 0, ctx->stream, Xi, Xo, b.m, nocc, pi, po,
                       state + 2 * p, state + 2 * (p ^ 1), state + 4, p ^ 1, tol, skip);
}

}  // namespace

int zq_sp2_enqueue(zq_ctx *ctx, const double *dF, int n, int nocc, double tol, int nsteps, const int *skip, int *state6,
                   const double **dX, long long *stride, int *ld, int track) {
  if (nsteps < 1 || nsteps > kSp2MaxSteps) return zq_fail(ctx, ZQ_ERR_INVALID, "SP2 step budget out of range");
  if (track < 0 || track > 2) return zq_fail(ctx, ZQ_ERR_INVALID, "SP2 fold tracking mode out of range");
  Sp2Bufs b;
  if (int rc = sp2_bufs(ctx, n, b)) return rc;
  zq_prof_begin(ctx, "scf_sp2");
  sp2_start(ctx, b, dF, n, nocc, state6, skip, track);
  double *est = track ? b.est : nullptr;
  // about kSp2TilesPerGroup tiles per workgroup and step: fewer workgroups hand each step over with fewer
  // release fences, at the cost of more tiles per workgroup on the step's critical path
  const int groups = std::max(1, (b.ntiles + kSp2TilesPerGroup - 1) / kSp2TilesPerGroup);
  if (b.m <= 256)
    hipLaunchKernelGGL(k_sp2_run<true>, dim3(groups), dim3(sp2_threads(true)), 0, ctx->stream, b.X, b.mm, b.m, nocc, b.part, b.sync,
                       state6, nsteps, tol, skip, est);
  else
    hipLaunchKernelGGL(k_sp2_run<false>, dim3(std::min(groups, 1024)), dim3(64), 0, ctx->stream, b.X, b.mm, b.m, nocc,
                       b.part, b.sync, state6, nsteps, tol, skip, est);
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, "k_sp2_run");
  *dX = b.X;
  *stride = (long long)b.mm;
  *ld = b.m;
  return ZQ_OK;
}

int zq_sp2_tracked_stats(zq_ctx *ctx, int out[3]) {
  auto *est = (const double *)zq_scratch(ctx, "sp2_est", kEstSize * sizeof(double));
  if (!est) return zq_fail(ctx, ZQ_ERR_OOM, "SP2 buffers");
  double h[3] = {0.0, 0.0, 0.0};
  ZQ_TRY(ctx, hipMemcpyAsync(h, est + kEstRuns, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < 3; ++k) out[k] = int(h[k]);
  return ZQ_OK;
}

int zq_sp2_density(zq_ctx *ctx, const double *dF, int n, int nocc, double tol, int max_steps, int *steps_hint,
                   const double **dD, int *ld, int *steps) {
  Sp2Bufs b;
  if (int rc = sp2_bufs(ctx, n, b)) return rc;
  auto *state = (int *)zq_scratch(ctx, "sp2_state", 6 * sizeof(int));
  if (!state) return zq_fail(ctx, ZQ_ERR_OOM, "SP2 buffers");
  zq_prof_begin(ctx, "scf_sp2");
  sp2_start(ctx, b, dF, n, nocc, state, nullptr, 0);
  int launched = 0, h[2] = {0, 0}, fin = 0;
  int next_check = std::max(4, std::min(max_steps, *steps_hint + 3));
  while (true) {
    while (launched < next_check) sp2_step(ctx, b, launched++, nocc, tol, state, nullptr);
    ZQ_LAUNCH_CHECK(ctx, "k_sp2_step");
    ZQ_TRY(ctx, hipMemcpyAsync(h, state + 2 * (launched & 1), 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (h[0] >= 2 || launched >= max_steps) break;
    next_check = std::min(max_steps, launched + 4);
  }
  ZQ_TRY(ctx, hipMemcpyAsync(&fin, state + 4, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  zq_prof_end(ctx);
  *dD = b.X + (fin & 1) * b.mm;
  *ld = b.m;
  *steps = h[1];
  if (h[0] < 2) return ZQ_ERR_NOT_CONVERGED;
  *steps_hint = h[1];
  return ZQ_OK;
}
