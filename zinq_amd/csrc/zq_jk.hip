// J/K Fock build over the resident 8-fold packed ERIs: the getFock contraction of
// src/hartree_fock.zig:686-735 (non-DFT branch),
//   F[k,l] = H[k,l] + sum_ij P[i,j] (g[i,k,j,l] - x g[i,j,k,l]),  F <- (F + F^T)/2,
// with J_kl = sum_ij P_ij (ij|kl) and K_kl = sum_ij P_ij (ik|jl) in chemist notation.
//
// Each unique integral v = (pq|rs) of degeneracy deg (number of distinct index permutations) is read
// once; with w = deg/8 it adds to an unsymmetrised accumulator G
//   G[p,q] += 2w P[r,s] v,  G[r,s] += 2w P[p,q] v,
//   G[q,s] -= x w P[p,r] v, G[p,s] -= x w P[q,r] v, G[q,r] -= x w P[p,s] v, G[p,r] -= x w P[q,s] v,
// and F = H + G + G^T, which is exactly symmetric (so the reference's final symmetrisation is implied).
//
// Work decomposition: one workgroup per (bra block pair, run of <= CH ket blocks); lane = bra pair
// (p,q) of the 8x8 bra block; the quartets stream through registers (coalesced 512-byte rows) and the six
// updates accumulate there, reduced across lanes by DPP moves and permlane swaps (no LDS).  Partial
// results are written as 8x8 "records" with a fixed target block.  A second kernel sums, for each
// output block, its records in a fixed order (a host-built CSR list): no atomics, bitwise reproducible.

#include <algorithm>
#include <string>
#include <vector>

#include "zq_internal.h"

namespace {

// Lane exchanges without LDS: DPP moves inside a 16-lane row and the gfx950 permlane swaps across rows.
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double dpp_xor1(double v) { return dpp_move<0xB1>(v); }                 // quad_perm 1 0 3 2
__device__ inline double dpp_xor2(double v) { return dpp_move<0x4E>(v); }                 // quad_perm 2 3 0 1
__device__ inline double dpp_xor4(double v) { return dpp_move<0x1B>(dpp_move<0x141>(v)); }  // half-row mirror, then ^3
__device__ inline double dpp_xor8(double v) { return dpp_move<0x128>(v); }                // row rotate by 8

// Lanes < 32 get x summed with lane ^ 32, lanes >= 32 get y summed likewise (v_permlane32_swap).
__device__ inline double perm32_sum(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane32_swap(unsigned(__double2loint(x)), unsigned(__double2loint(y)), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(unsigned(__double2hiint(x)), unsigned(__double2hiint(y)), false, false);
  return __hiloint2double(int(hi[0]), int(lo[0])) + __hiloint2double(int(hi[1]), int(lo[1]));
}
// Lanes with bit 4 clear get x summed with lane ^ 16, the others y (v_permlane16_swap).
__device__ inline double perm16_sum(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane16_swap(unsigned(__double2loint(x)), unsigned(__double2loint(y)), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(unsigned(__double2hiint(x)), unsigned(__double2hiint(y)), false, false);
  return __hiloint2double(int(hi[0]), int(lo[0])) + __hiloint2double(int(hi[1]), int(lo[1]));
}

// Sum of each of 16 values over the 64 lanes; lane l ends with the sum of value l >> 2 (reduce-scatter
// over lane bits 5, 4, 3, 2, then the two low bits).
__device__ inline double sum16_scatter(const double (&v)[16], int lane) {
  double u[8], w[4], z[2];
#pragma unroll
  for (int i = 0; i < 8; ++i) u[i] = perm32_sum(v[i], v[i + 8]);
#pragma unroll
  for (int i = 0; i < 4; ++i) w[i] = perm16_sum(u[i], u[i + 4]);
  const bool b3 = lane & 8, b2 = lane & 4;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double send = b3 ? w[i] : w[i + 2], keep = b3 ? w[i + 2] : w[i];
    z[i] = keep + dpp_xor8(send);
  }
  const double send = b2 ? z[0] : z[1], keep = b2 ? z[1] : z[0];
  double t = keep + dpp_xor4(send);
  t += dpp_xor2(t);
  t += dpp_xor1(t);
  return t;
}

// Sums of a[0], a[1] over lane bits 0-2 (q); lane l ends with the sum of a[l & 1].
__device__ inline double sum2_over_q(const double (&a)[2], int lane) {
  const bool b0 = lane & 1;
  const double send = b0 ? a[0] : a[1], keep = b0 ? a[1] : a[0];
  double t = keep + dpp_xor1(send);
  t += dpp_xor2(t);
  t += dpp_xor4(t);
  return t;
}

// Sums of a[0], a[1] over lane bits 3-5 (p); lane l ends with the sum of a[l >> 5].
__device__ inline double sum2_over_p(const double (&a)[2]) {
  double t = perm32_sum(a[0], a[1]);
  t = perm16_sum(t, t);
  t += dpp_xor8(t);
  return t;
}

// Reduce-scatter of 8 values over lane bits 2, 1, 0: lane l ends with the sum of v[l & 7] over the 8
// lanes that share l >> 3.
__device__ inline double sum8_scatter_low(const double (&v)[8], int lane) {
  const bool b2 = lane & 4, b1 = lane & 2, b0 = lane & 1;
  double a[4], b[2];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double send = b2 ? v[i] : v[i + 4], keep = b2 ? v[i + 4] : v[i];
    a[i] = keep + dpp_xor4(send);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double send = b1 ? a[i] : a[i + 2], keep = b1 ? a[i + 2] : a[i];
    b[i] = keep + dpp_xor2(send);
  }
  const double send = b0 ? b[0] : b[1], keep = b0 ? b[1] : b[0];
  return keep + dpp_xor1(send);
}

// Reduce-scatter of 8 values over lane bits 5, 4, 3: lane l ends with the sum of v[l >> 3] over the 8
// lanes that share l & 7.
__device__ inline double sum8_scatter_high(const double (&v)[8], int lane) {
  double a[4], b[2];
#pragma unroll
  for (int i = 0; i < 4; ++i) a[i] = perm32_sum(v[i], v[i + 4]);
#pragma unroll
  for (int i = 0; i < 2; ++i) b[i] = perm16_sum(a[i], a[i + 2]);
  const bool b3 = lane & 8;
  const double send = b3 ? b[0] : b[1], keep = b3 ? b[1] : b[0];
  return keep + dpp_xor8(send);
}

// One workgroup (4 waves) per work item; wave w handles ket rows r = 2w, 2w+1 of every quartet for all
// 64 bra pairs (lane = bra pair (p,q)), loading its two 512-byte rows straight into registers.  While a
// quartet is being contracted, the next quartet's rows and the per-lane density values it needs are
// already in flight (they are issued before any load the current contraction waits on, so the in-order
// vmcnt wait never stalls on them); density values common to the wave come in through scalar loads.
// Cross-wave sums (K[p,s], K[q,s] at the end of a bs run, J[p,q] at the end of the item) go through
// LDS in a fixed order.
// GROUP: the item's molecule selects the quartets, P and its leading dimension from mols[], and the
// item is skipped once done[molecule] is set.
// MULTI: blockIdx.x is the density d (P + d * pstride, records + d * rstride) and blockIdx.y the item,
// so the densities of one item run back to back and share its quartet reads through the caches.
template <class E, bool GROUP, bool MULTI = false>
__global__ __launch_bounds__(256) void zq_jk_records(const E *__restrict__ eri, const JKItem *__restrict__ items,
                                                     int nitems, const double *__restrict__ P, int npad, double xexch,
                                                     double *__restrict__ rec, const int *__restrict__ skip,
                                                     const JKMol *__restrict__ mols, const int *__restrict__ done,
                                                     long long pstride = 0, long long rstride = 0) {
  __shared__ double red[2][4][2][64];  // two buffers, so each cross-wave sum needs one barrier
  const int item_id = MULTI ? blockIdx.y : blockIdx.x;
  if (MULTI) {
    P += blockIdx.x * pstride;
    rec += blockIdx.x * rstride;
  }
  if (item_id >= nitems || (skip && *skip)) return;
  if (GROUP) {
    const int m = items[item_id].mol;
    if (done && done[m]) return;
    eri += mols[m].eri_off;
    P += mols[m].p_off;
    npad = mols[m].ld;
  }
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int Ib = items[item_id].Ib, bp = items[item_id].bp, bq = items[item_id].bq;
  const int k0 = items[item_id].k0, k1 = items[item_id].k1, rec0 = items[item_id].rec0;

  const int p = lane >> 3, q = lane & 7;
  const int Pg = bp * 8 + p, Qg = bq * 8 + q;
  // fI halves the bra pairs counted twice (p == q on a diagonal bra block).  It is a power of two and
  // constant per lane, so it is folded into this lane's density values (exactly) rather than into every
  // integral; Jbra, which takes no such value, is scaled once at the end.
  const double fI = (bp == bq && p == q) ? 0.5 : 1.0;
  const double *Prow = P + size_t(Pg) * npad, *Qrow = P + size_t(Qg) * npad;
  const double Ppq = Prow[Qg] * fI;

  int bs = 0, kk = k0;
  for (;; ++bs) {
    const int len = zq_run_len(bp, bq, bs);
    if (kk < len) break;
    kk -= len;
  }
  int br = bs + kk;

  double *out = rec + size_t(rec0) * 64;
  int rb = 0;  // buffer of red for the next cross-wave sum
  // this wave's two rows of the item's first quartet: element (r*8+s)*64 + lane, r = 2w + rr
  const E *src = eri + (size_t(Ib) * (Ib + 1) / 2 + size_t(k0)) * ZQ_QUARTET + (2 * wave) * 512 + lane;
  double Jbra = 0.0;
  // quartets in flight ahead of the contraction, kept in their stored type (4-byte ones leave room for 4)
  constexpr int D = sizeof(E) == 4 ? 4 : 2;
  double Kps[8], Kqs[8], Pp_s[8], Pq_s[8], cur[16];
  E pf[D][16];
  double Pr_cur[4], Pr_nxt[4];  // Prow[br*8+r], Qrow[br*8+r] for the wave's two rows
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    Kps[s] = Kqs[s] = 0.0;
    Pp_s[s] = Prow[bs * 8 + s] * fI;
    Pq_s[s] = Qrow[bs * 8 + s] * fI;
  }
#pragma unroll
  for (int rr = 0; rr < 2; ++rr) {
    Pr_cur[2 * rr] = Prow[br * 8 + 2 * wave + rr] * fI;
    Pr_cur[2 * rr + 1] = Qrow[br * 8 + 2 * wave + rr] * fI;
  }
  // the quartet stream is read once per build: non-temporal loads, D quartets ahead of the contraction
#pragma unroll
  for (int j = 0; j < 16; ++j) cur[j] = double(__builtin_nontemporal_load(src + j * 64));
#pragma unroll
  for (int d = 0; d + 1 < D; ++d)
    if (k0 + 1 + d < k1) {
#pragma unroll
      for (int j = 0; j < 16; ++j) pf[d][j] = __builtin_nontemporal_load(src + size_t(1 + d) * ZQ_QUARTET + j * 64);
    }

  for (int k = k0; k < k1; ++k) {
    if (k + D < k1) {
      const E *nxd = src + size_t(k + D - k0) * ZQ_QUARTET;
#pragma unroll
      for (int j = 0; j < 16; ++j) pf[D - 1][j] = __builtin_nontemporal_load(nxd + j * 64);
    }
    const bool run_end = (br + 1 > (bs <= bq ? bp : bp - 1));
    int nbs = bs, nbr = br + 1;
    if (run_end) {
      ++nbs;
      while (nbs <= bp && zq_run_len(bp, bq, nbs) <= 0) ++nbs;
      nbr = nbs;
    }
    if (k + 1 < k1) {
#pragma unroll
      for (int rr = 0; rr < 2; ++rr) {
        Pr_nxt[2 * rr] = Prow[nbr * 8 + 2 * wave + rr] * fI;
        Pr_nxt[2 * rr + 1] = Qrow[nbr * 8 + 2 * wave + rr] * fI;
      }
      // quartet k + 1 is already in flight (pf[0])
    }
    const int Jb = zq_bpair(br, bs);
    const bool diagJ = (br == bs), diagQ = (Jb == Ib);
    if (diagJ || diagQ) {  // (uniform, rare) halve the ket pairs r == s and the bra-ket diagonal
#pragma unroll
      for (int rr = 0; rr < 2; ++rr)
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const int r = 2 * wave + rr;
          double w = 1.0;
          if (diagJ && r == s) w = 0.5;
          if (diagQ && lane == r * 8 + s) w *= 0.5;
          cur[rr * 8 + s] *= w;
        }
    }
    double jt[16], accpr[2], accqr[2];
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
      const int r = 2 * wave + rr;
      const double Pp_r = Pr_cur[2 * rr], Pq_r = Pr_cur[2 * rr + 1];
      const double *Prs = P + size_t(br * 8 + r) * npad + bs * 8;  // wave-uniform
      double apr = 0.0, aqr = 0.0;
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const double a = cur[rr * 8 + s];
        Jbra += Prs[s] * a;
        jt[rr * 8 + s] = Ppq * a;
        Kqs[s] += Pp_r * a;
        Kps[s] += Pq_r * a;
        aqr += Pp_s[s] * a;
        apr += Pq_s[s] * a;
      }
      accpr[rr] = apr;
      accqr[rr] = aqr;
    }
    // J of the wave's 16 ket pairs (r, s), over the 64 bra pairs: lane l holds pair l >> 2
    const double tj = sum16_scatter(jt, lane);
    if ((lane & 3) == 0) out[(2 * wave + (lane >> 5)) * 8 + ((lane >> 2) & 7)] = 2.0 * tj;
    // K[p, br*8+r] over q (lanes with q = rr write row rr) and K[q, br*8+r] over p (p = 4 rr writes row rr)
    const double kp = sum2_over_q(accpr, lane), kq = sum2_over_p(accqr);
    if (q < 2) out[64 + p * 8 + 2 * wave + q] = -xexch * kp;
    if ((p & 3) == 0) out[128 + q * 8 + 2 * wave + (p >> 2)] = -xexch * kq;
    out += 192;

    if (run_end || k == k1 - 1) {
      // buffer rb was last read by wave 0 before the previous cross-wave barrier, so it is free
      red[rb][wave][0][lane] = sum8_scatter_low(Kps, lane);   // K[p, bs*8+s], s = q
      red[rb][wave][1][lane] = sum8_scatter_high(Kqs, lane);  // K[q, bs*8+s], s = p
      __syncthreads();
      if (wave == 0) {
        const double a = ((red[rb][0][0][lane] + red[rb][1][0][lane]) + red[rb][2][0][lane]) + red[rb][3][0][lane];
        const double b = ((red[rb][0][1][lane] + red[rb][1][1][lane]) + red[rb][2][1][lane]) + red[rb][3][1][lane];
        out[p * 8 + q] = -xexch * a;
        out[64 + q * 8 + p] = -xexch * b;
      }
      rb ^= 1;
      out += 128;
#pragma unroll
      for (int s = 0; s < 8; ++s) Kps[s] = Kqs[s] = 0.0;
      if (run_end && k + 1 < k1) {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          Pp_s[s] = Prow[nbs * 8 + s] * fI;
          Pq_s[s] = Qrow[nbs * 8 + s] * fI;
        }
      }
    }
    bs = nbs, br = nbr;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      cur[j] = double(pf[0][j]);
#pragma unroll
      for (int d = 0; d + 1 < D; ++d) pf[d][j] = pf[d + 1][j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) Pr_cur[j] = Pr_nxt[j];
  }
  red[rb][wave][0][lane] = Jbra * fI;
  __syncthreads();
  if (wave == 0)
    out[lane] = 2.0 * (((red[rb][0][0][lane] + red[rb][1][0][lane]) + red[rb][2][0][lane]) + red[rb][3][0][lane]);
}

// F block (x >= y) = H + sum of records aimed at (x,y) + transposed records aimed at (y,x).  kRedWaves
// waves take contiguous shares of the block's record list with 8 loads in flight each; the shares are
// added in wave order, so the result is bitwise reproducible.
// GROUP: blk[blockIdx.x] = (molecule, its output block); H, F, n and the leading dimension come from
// mols[].  Otherwise H and F are n x n with leading dimension ldf and blockIdx.x is the output block.
// MULTI: blockIdx.y is the density (records + d * rstride, F + d * fstride); H may be null (base 0).
// epart (optional, plain form): epart[block] = this block's share of 1/2 sum P (Hc + F), the electronic
// energy of getEnergy (src/hartree_fock.zig:643-655), from the padded density Pe (ld pld) and the core
// Hamiltonian Hc (ld ldf); blocks off the diagonal count twice, as their mirrors are not visited.
constexpr int kRedWaves = 16;

template <bool GROUP, bool MULTI = false>
__global__ __launch_bounds__(64 * kRedWaves) void zq_jk_reduce(const double *__restrict__ rec, const int *__restrict__ csr_ptr,
                                                    const int *__restrict__ csr_idx, const double *H, int n, double *F,
                                                    const int *__restrict__ skip, int ldf, const JKMol *__restrict__ mols,
                                                    const int2 *__restrict__ blk, const int *__restrict__ done,
                                                    long long rstride = 0, long long fstride = 0,
                                                    const double *__restrict__ Pe = nullptr, int pld = 0,
                                                    const double *__restrict__ Hc = nullptr, double *__restrict__ epart = nullptr) {
  __shared__ double part[kRedWaves][64];
  if (skip && *skip) return;
  if (MULTI) {
    rec += blockIdx.y * rstride;
    F += blockIdx.y * fstride;
  }
  int o = blockIdx.x;
  if (GROUP) {
    const int2 b = blk[blockIdx.x];
    if (done && done[b.x]) return;
    o = b.y;
    H += mols[b.x].h_off;
    F += mols[b.x].f_off;
    n = mols[b.x].n;
    ldf = mols[b.x].ld;
  }
  const int g = blockIdx.x;  // row of the CSR (global over a group)
  int x, y;
  zq_unpair(o, x, y);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, i = lane >> 3, j = lane & 7;
  const int e0 = csr_ptr[g], e1 = csr_ptr[g + 1], len = e1 - e0;
  const int a0 = e0 + (len * wave) / kRedWaves, a1 = e0 + (len * (wave + 1)) / kRedWaves;
  double acc = 0.0;
  int e = a0;
  for (; e + 8 <= a1; e += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int code = csr_idx[e + u];
      const double *r = rec + size_t(code >> 1) * 64;
      v[u] = (code & 1) ? r[j * 8 + i] : r[lane];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += v[u];
  }
  for (; e < a1; ++e) {
    const int code = csr_idx[e];
    const double *r = rec + size_t(code >> 1) * 64;
    acc += (code & 1) ? r[j * 8 + i] : r[lane];
  }
  part[wave][lane] = acc;
  __syncthreads();
  if (wave != 0) return;
  const int gi = x * 8 + i, gj = y * 8 + j;
  const double h = (H && gi < n && gj < n) ? H[size_t(gi) * ldf + gj] : 0.0;
  double sum = part[0][lane];
#pragma unroll
  for (int k = 1; k < kRedWaves; ++k) sum += part[k][lane];
  const double f = h + sum;
  // on diagonal blocks lane (i,j) and lane (j,i) sum the same terms in different orders: let only
  // i >= j write, so both triangles hold one value
  if (gi < n && gj < n && (x != y || i >= j)) {
    F[size_t(gi) * ldf + gj] = f;
    F[size_t(gj) * ldf + gi] = f;
  }
  if (!GROUP && !MULTI && epart) {
    const double w = x != y ? 1.0 : (i > j ? 1.0 : (i == j ? 0.5 : 0.0));
    double c = (gi < n && gj < n) ? w * Pe[size_t(gi) * pld + gj] * (Hc[size_t(gi) * ldf + gj] + f) : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (lane == 0) epart[blockIdx.x] = c;
  }
}

}  // namespace

// Host-side work plan: work items, the exact record sequence each emits, and the per-output-block CSR.
void zq_jk_plan(int Nb, size_t nq, int chunk, JKPlan &plan) {
  const int CH = chunk > 0 ? chunk : int(std::max<size_t>(1, std::min<size_t>(16, nq / 3072)));
  std::vector<JKItem> &items = plan.items;
  items.clear();
  std::vector<std::pair<int, int>> tgt;
  for (int Ib = 0; Ib < Nb; ++Ib) {
    int bp, bq;
    zq_unpair(Ib, bp, bq);
    const int total = Ib + 1;
    for (int k0 = 0; k0 < total; k0 += CH) {
      const int k1 = std::min(total, k0 + CH);
      items.push_back({Ib, bp, bq, k0, k1, int(tgt.size()), 0});
      int bs = 0, kk = k0;
      for (;; ++bs) {
        const int len = zq_run_len(bp, bq, bs);
        if (kk < len) break;
        kk -= len;
      }
      int br = bs + kk;
      for (int k = k0; k < k1; ++k) {
        tgt.push_back({br, bs});
        tgt.push_back({bp, br});
        tgt.push_back({bq, br});
        const bool run_end = (br + 1 > (bs <= bq ? bp : bp - 1));
        if (run_end || k == k1 - 1) {
          tgt.push_back({bp, bs});
          tgt.push_back({bq, bs});
        }
        if (run_end) {
          ++bs;
          while (bs <= bp && zq_run_len(bp, bq, bs) <= 0) ++bs;
          br = bs;
        } else {
          ++br;
        }
      }
      tgt.push_back({bp, bq});
    }
  }
  // CSR over output blocks x >= y
  std::vector<std::vector<int>> lists(Nb);
  for (size_t r = 0; r < tgt.size(); ++r) {
    const int tx = tgt[r].first, ty = tgt[r].second;
    if (tx >= ty) lists[zq_bpair(tx, ty)].push_back(int(r) << 1);
  }
  for (size_t r = 0; r < tgt.size(); ++r) {
    const int tx = tgt[r].first, ty = tgt[r].second;
    if (ty >= tx) lists[zq_bpair(ty, tx)].push_back((int(r) << 1) | 1);
  }
  plan.ptr.assign(Nb + 1, 0);
  plan.idx.clear();
  for (int o = 0; o < Nb; ++o) {
    plan.ptr[o + 1] = plan.ptr[o] + int(lists[o].size());
    plan.idx.insert(plan.idx.end(), lists[o].begin(), lists[o].end());
  }
  plan.nrecords = tgt.size();
}

void zq_jk_plan_free(JKPlanDev &d) {
  for (void *p : {(void *)d.items, (void *)d.ptr, (void *)d.idx})
    if (p) (void)hipFree(p);
  d = JKPlanDev{};
}

int zq_jk_plan_upload(int Nb, size_t nquartets, hipStream_t st, JKPlanDev &d, std::string &err) {
  JKPlan plan;
  zq_jk_plan(Nb, nquartets, 0, plan);
  if (plan.nrecords >= (size_t(1) << 30)) {
    err = "too many J/K records";
    return ZQ_ERR_INVALID;
  }
  const std::vector<JKItem> &items = plan.items;
  const std::vector<int> &ptr = plan.ptr, &idx = plan.idx;
  JKPlanDev out;
  out.nitems = int(items.size());
  out.nrecords = plan.nrecords;
  // uploads on st, never the null stream: st is non-blocking, and null-stream work is not ordered
  // before its kernels
  hipError_t e = hipMalloc(&out.items, items.size() * sizeof(JKItem));
  if (e == hipSuccess) e = hipMalloc(&out.ptr, ptr.size() * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&out.idx, std::max<size_t>(1, idx.size()) * sizeof(int));
  if (e == hipSuccess) e = hipMemcpyAsync(out.items, items.data(), items.size() * sizeof(JKItem), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(out.ptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !idx.empty())
    e = hipMemcpyAsync(out.idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // the plan's host vectors go out of scope
  if (e != hipSuccess) {
    err = std::string("J/K plan upload: ") + hipGetErrorString(e);
    zq_jk_plan_free(out);
    return e == hipErrorOutOfMemory ? ZQ_ERR_OOM : ZQ_ERR_LAUNCH;
  }
  d = out;
  return ZQ_OK;
}

// The context's records, P (npad x npad), H and F (n x n) for its current size.  Each grows when too
// small and is otherwise kept: a hipFree waits for every stream on the device, which under concurrent
// requests would stall an attach behind the other slots' work.  P is zeroed on the context's stream.
int zq_jk_buffers(zq_ctx *ctx) {
  const size_t npad2 = size_t(ctx->npad) * ctx->npad, nn = size_t(ctx->n) * ctx->n;
  struct Want {
    double **p;
    size_t *cap, need;
  };
  for (const Want &w : {Want{&ctx->d_records, &ctx->cap_records, ctx->nrecords * 64}, Want{&ctx->d_P, &ctx->cap_P, npad2},
                        Want{&ctx->d_H, &ctx->cap_H, nn}, Want{&ctx->d_F, &ctx->cap_F, nn}}) {
    if (*w.cap >= w.need && *w.p) continue;
    if (*w.p) (void)hipFree(*w.p);
    *w.p = nullptr;
    *w.cap = 0;
    ++ctx->scratch_gen;  // captured SCF iterations hold the old address
    ZQ_TRY(ctx, hipMalloc(w.p, std::max<size_t>(1, w.need) * sizeof(double)));
    *w.cap = w.need;
  }
  ZQ_TRY(ctx, hipMemsetAsync(ctx->d_P, 0, npad2 * sizeof(double), ctx->stream));
  return ZQ_OK;
}

// zq_eri_load: the context's own plan for its ERIs, then its buffers.
int zq_build_jk_plan(zq_ctx *ctx) {
  JKPlanDev d;
  std::string err;
  if (int rc = zq_jk_plan_upload(ctx->Nb, ctx->nquartets, ctx->stream, d, err)) return zq_fail(ctx, rc, err);
  ctx->d_items = d.items, ctx->d_csr_ptr = d.ptr, ctx->d_csr_idx = d.idx;
  ctx->nitems = d.nitems, ctx->nrecords = d.nrecords;
  ctx->plan_owned = true;
  return zq_jk_buffers(ctx);
}

// Device-resident J/K: P (npad x npad, zero padded) and H (n x n) already on the device; F (n x n) out.
// With skip != nullptr both kernels return at once while *skip != 0 (device-side early exit).  With
// f32 the records kernel streams the FP32 copy of the quartets (zq_eri32); products and sums stay FP64.
// dH may equal dF: every output block reads its base before its own wave writes it, and the blocks
// above the diagonal are only written (mirror), never read.  epart (Nb doubles, optional): the blocks'
// shares of the electronic energy 1/2 sum P (H + F) with the context's P and core H.
int zq_fock_jk_device(zq_ctx *ctx, const double *dP, const double *dH, double x_exch, double *dF, const int *skip,
                      bool f32, double *epart) {
  if (f32 && !ctx->eri32) return zq_fail(ctx, ZQ_ERR_STATE, "zq_fock_jk_device: no FP32 ERI copy");
  zq_prof_begin(ctx, f32 ? "jk_records32" : "jk_records");
  if (f32)
    hipLaunchKernelGGL((zq_jk_records<float, false>), dim3(unsigned(ctx->nitems)), dim3(256), 0, ctx->stream,
                       ctx->eri32, ctx->d_items, ctx->nitems, dP, ctx->npad, x_exch, ctx->d_records, skip, nullptr,
                       nullptr);
  else
    hipLaunchKernelGGL((zq_jk_records<double, false>), dim3(unsigned(ctx->nitems)), dim3(256), 0, ctx->stream,
                       ctx->eri, ctx->d_items, ctx->nitems, dP, ctx->npad, x_exch, ctx->d_records, skip, nullptr,
                       nullptr);
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, "zq_jk_records");
  zq_prof_begin(ctx, "jk_reduce");
  hipLaunchKernelGGL(zq_jk_reduce<false>, dim3(unsigned(ctx->Nb)), dim3(64 * kRedWaves), 0, ctx->stream, ctx->d_records,
                     ctx->d_csr_ptr, ctx->d_csr_idx, dH, ctx->n, dF, skip, ctx->n, nullptr, nullptr, nullptr, 0LL, 0LL,
                     ctx->d_P, ctx->npad, ctx->d_H, epart);
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, "zq_jk_reduce");
  return ZQ_OK;
}

hipError_t zq_jk_group(hipStream_t stream, const double *eri, const JKItem *items, int nitems, const double *P,
                       double x_exch, double *rec, const JKMol *mols, const int *done, const int *csr_ptr,
                       const int *csr_idx, const int2 *blk, int nblk, const double *H, double *F) {
  hipLaunchKernelGGL((zq_jk_records<double, true>), dim3(unsigned(nitems)), dim3(256), 0, stream, eri, items, nitems, P,
                     0, x_exch, rec, nullptr, mols, done);
  hipLaunchKernelGGL(zq_jk_reduce<true>, dim3(unsigned(nblk)), dim3(64 * kRedWaves), 0, stream, rec, csr_ptr, csr_idx, H, 0, F,
                     nullptr, 0, mols, blk, done);
  return hipGetLastError();
}

namespace {
__global__ void zq_narrow(const double *__restrict__ x, size_t n, float *__restrict__ y) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    y[i] = float(x[i]);
}
}  // namespace

void zq_narrow_eri(hipStream_t stream, const double *eri, size_t cnt, float *out) {
  hipLaunchKernelGGL(zq_narrow, dim3(4096), dim3(256), 0, stream, eri, cnt, out);
}

// FP32 copy of the resident quartets (round to nearest), made once per ERI load.
int zq_eri32(zq_ctx *ctx) {
  if (ctx->eri32) return ZQ_OK;
  if (!ctx->eri) return zq_fail(ctx, ZQ_ERR_STATE, "zq_eri32: no ERIs loaded");
  const size_t cnt = ctx->nquartets * ZQ_QUARTET;
  ZQ_TRY(ctx, hipMalloc(&ctx->eri32, cnt * sizeof(float)));
  ctx->eri32_owned = true;
  ++ctx->scratch_gen;  // a new address for captured SCF iterations (the residency key stays)
  zq_narrow_eri(ctx->stream, ctx->eri, cnt, ctx->eri32);
  ZQ_LAUNCH_CHECK(ctx, "zq_narrow");
  return ZQ_OK;
}

extern "C" int zq_fock_jk_multi(zq_ctx *ctx, const double *H, const double *P, size_t ndens, double x_exch,
                                double *F_out) {
  ZqGate _gate;
  if (!ctx || !P || !F_out || ndens == 0 || ndens > 65535)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_fock_jk_multi: bad arguments");
  if (!ctx->eri) return zq_fail(ctx, ZQ_ERR_STATE, "zq_fock_jk_multi: no ERIs loaded");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  const int n = ctx->n, np = ctx->npad;
  const size_t nn = size_t(n) * n, pstride = size_t(np) * np, rstride = ctx->nrecords * 64;
  auto *dP = (double *)zq_scratch(ctx, "jkm_P", ndens * pstride * sizeof(double), nullptr, false);
  auto *dR = (double *)zq_scratch(ctx, "jkm_rec", ndens * rstride * sizeof(double), nullptr, false);
  auto *dF = (double *)zq_scratch(ctx, "jkm_F", ndens * nn * sizeof(double), nullptr, false);
  if (!dP || !dR || !dF) return zq_fail(ctx, ZQ_ERR_OOM, "zq_fock_jk_multi: scratch");
  // each density into its own zero-padded npad x npad slot
  ZQ_TRY(ctx, hipMemsetAsync(dP, 0, ndens * pstride * sizeof(double), ctx->stream));
  for (size_t d = 0; d < ndens; ++d)
    ZQ_TRY(ctx, hipMemcpy2DAsync(dP + d * pstride, np * sizeof(double), P + d * nn, n * sizeof(double),
                                 n * sizeof(double), n, hipMemcpyHostToDevice, ctx->stream));
  if (H) ZQ_TRY(ctx, hipMemcpyAsync(ctx->d_H, H, nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  zq_prof_begin(ctx, "jk_records");
  hipLaunchKernelGGL((zq_jk_records<double, false, true>), dim3(unsigned(ndens), unsigned(ctx->nitems)), dim3(256), 0,
                     ctx->stream, ctx->eri, ctx->d_items, ctx->nitems, dP, np, x_exch, dR, nullptr, nullptr, nullptr,
                     (long long)pstride, (long long)rstride);
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, "zq_jk_records (multi)");
  zq_prof_begin(ctx, "jk_reduce");
  hipLaunchKernelGGL((zq_jk_reduce<false, true>), dim3(unsigned(ctx->Nb), unsigned(ndens)), dim3(64 * kRedWaves), 0, ctx->stream, dR,
                     ctx->d_csr_ptr, ctx->d_csr_idx, H ? ctx->d_H : nullptr, n, dF, nullptr, n, nullptr, nullptr, nullptr,
                     (long long)rstride, (long long)nn);
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, "zq_jk_reduce (multi)");
  ZQ_TRY(ctx, hipMemcpyAsync(F_out, dF, ndens * nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ZQ_OK;
}

extern "C" int zq_fock_jk(zq_ctx *ctx, const double *H, const double *P, double x_exch, double *F_out) {
  ZqGate _gate;
  if (!ctx || !H || !P || !F_out) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_fock_jk: null argument");
  if (!ctx->eri) return zq_fail(ctx, ZQ_ERR_STATE, "zq_fock_jk: no ERIs loaded");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  const int n = ctx->n, np = ctx->npad;
  ZQ_TRY(ctx, hipMemcpy2DAsync(ctx->d_P, np * sizeof(double), P, n * sizeof(double), n * sizeof(double), n,
                               hipMemcpyHostToDevice, ctx->stream));
  ZQ_TRY(ctx, hipMemcpyAsync(ctx->d_H, H, size_t(n) * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  int rc = zq_fock_jk_device(ctx, ctx->d_P, ctx->d_H, x_exch, ctx->d_F);
  if (rc) return rc;
  ZQ_TRY(ctx, hipMemcpyAsync(F_out, ctx->d_F, size_t(n) * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ZQ_OK;
}

extern "C" int zq_grad_eri(zq_ctx *ctx, const double *dg_packed, size_t n, size_t ncoord, const double *P, double x_exch,
                           double *out) {
  ZqGate _gate;
  if (!ctx || !dg_packed || !P || !out || n == 0 || n > 4096 || ncoord == 0)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_grad_eri: bad arguments");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  // a private context on the same device holds each coordinate's derivative ERIs in turn, so ctx's own
  // resident ERIs, plan and buffers stay as they are
  zq_ctx *w = zq_create(ctx->device);
  if (!w) return zq_fail(ctx, ZQ_ERR_OOM, "zq_grad_eri: work context");
  const size_t np = n * (n + 1) / 2, nu = np * (np + 1) / 2;
  int rc = ZQ_OK;
  std::vector<double> part;
  for (size_t x = 0; x < ncoord && rc == ZQ_OK; ++x) {
    rc = zq_eri_load_packed(w, dg_packed + x * nu, n);
    if (rc) break;
    auto *epart = (double *)zq_scratch(w, "grad_epart", size_t(w->Nb) * sizeof(double), nullptr, false);
    if (!epart) {
      rc = zq_fail(w, ZQ_ERR_OOM, "zq_grad_eri: energy partials");
      break;
    }
    // 1/2 sum P (Hc + F) with Hc = 0 and F = G(P) of this set: the reduce's energy partials
    hipError_t e = hipMemsetAsync(w->d_H, 0, n * n * sizeof(double), w->stream);
    if (e == hipSuccess)
      e = hipMemcpy2DAsync(w->d_P, w->npad * sizeof(double), P, n * sizeof(double), n * sizeof(double), n,
                           hipMemcpyHostToDevice, w->stream);
    if (e != hipSuccess) {
      rc = zq_check(w, e, "zq_grad_eri: staging");
      break;
    }
    rc = zq_fock_jk_device(w, w->d_P, w->d_H, x_exch, w->d_F, nullptr, false, epart);
    if (rc) break;
    part.assign(w->Nb, 0.0);
    e = hipMemcpyAsync(part.data(), epart, part.size() * sizeof(double), hipMemcpyDeviceToHost, w->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) {
      rc = zq_check(w, e, "zq_grad_eri: read-back");
      break;
    }
    double sum = 0.0;
    for (double v : part) sum += v;  // output blocks in order
    out[x] = sum;
  }
  if (rc) zq_fail(ctx, rc, std::string("zq_grad_eri: ") + zq_last_error(w));
  zq_destroy(w);
  return rc;
}

extern "C" size_t zq_jk_record_count(size_t n) {
  if (n == 0 || n > 4096) return 0;
  const int nb = int((n + 7) / 8), Nb = nb * (nb + 1) / 2;
  JKPlan plan;
  zq_jk_plan(Nb, size_t(Nb) * (Nb + 1) / 2, 0, plan);
  return plan.nrecords;
}
