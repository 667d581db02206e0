// AO->MO four-index transform of the resident ERIs and the MP2 energy that consumes it.
//
// Reference: ao2mo_pppp (src/integral_transform.zig:37-79) and its occupied/virtual restriction
// ao2mo_oovv (:165-219) compute, as four quarter transforms in physicist order,
//   out[i,j,a,b] = sum C[mu,i] C[lam,j] C[nu,a] C[sig,b] g[mu,lam,nu,sig],  g[mu,lam,nu,sig] = (mu nu|lam sig).
// Here the same chain runs as four GEMM stages on the matrix cores:
//   Q1  X[K][i][nu]      = sum_mu  C1[mu,i]  (mu nu|K)         for every canonical ket pair K = (lam >= sig)
//   Q2  Y[i][j][sig][nu] = sum_lam C2[lam,j] X[(lam sig)][i][nu]   (X is symmetric in lam, sig)
//   Q3  W[i][j][sig][a]  = sum_nu  Y[i][j][sig][nu] C3[nu,a]
//   Q4  out[i][j][a][b]  = sum_sig W[i][j][sig][a] C4[sig,b]
// Q1 reads the 8x8 block quartets straight from HBM into LDS (no unpacked copy of the ERIs) and uses
// the ket-pair symmetry, so its executed flops are half of the dense chain's first quarter.
//
// Precisions of the quarter transforms:
//   FP64    v_mfma_f64_16x16x4f64, FP64 intermediates;
//   FP32    v_mfma_f32_16x16x4f32, FP32 intermediates;
//   FP32C   v_mfma_f32_16x16x4f32 products summed in FP32 over one K block (16 k; 8 mu in Q1), the
//           blocks summed in FP64 (a compensated sum of FP32 products); FP32 intermediates;
//   BF16X3  every operand split x = hi + lo (two bf16), three v_mfma_f32_16x16x32_bf16 products
//           hi*hi + (hi*lo + lo*hi) with the small products in a separate FP32 accumulator that is added
//           at the end; FP32 intermediates.
// The MP2 denominators and sums are FP64 in every mode.

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>


#include "zq_internal.h"

namespace {

constexpr int kMaxGridZ = 65535;  // largest gridDim.z of one launch

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16;
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and B[l/16][l%16] and receives
// D[4*j + l/16][l%16] in accumulator element j (measured on gfx950).
struct MF64 {
  typedef double T;
  typedef d4 acc;
  static constexpr int kAccRegs = 8;  // 32-bit registers of one accumulator tile
  static __device__ acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
  static __device__ acc mma(T a, T b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ int row(int lane, int j) { return 4 * j + (lane >> 4); }
  static __device__ void flush(acc &) {}
  static __device__ double val(const acc &c, int j) { return c[j]; }
};

// v_mfma_f32_16x16x4f32: same operand layout, D[4*(l/16) + j][l%16].
struct MF32 {
  typedef float T;
  typedef f4 acc;
  static constexpr int kAccRegs = 4;
  static __device__ acc zero() { return acc{0.f, 0.f, 0.f, 0.f}; }
  static __device__ acc mma(T a, T b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ int row(int lane, int j) { return 4 * (lane >> 4) + j; }
  static __device__ void flush(acc &) {}
  static __device__ double val(const acc &c, int j) { return c[j]; }
};

// FP32 products with FP64 sums: v_mfma_f32_16x16x4f32 accumulates one K block (16 k in the GEMMs, 8 mu
// in Q1) in FP32; flush() adds the block's partial sums into FP64 accumulators and restarts the FP32
// ones.  The FP32 rounding of the running sum then only spans one block instead of all of K.
struct MF32C {
  typedef float T;
  struct acc {
    f4 p;
    d4 s;
  };
  static constexpr int kAccRegs = 12;
  static __device__ acc zero() { return acc{f4{0.f, 0.f, 0.f, 0.f}, d4{0.0, 0.0, 0.0, 0.0}}; }
  static __device__ acc mma(T a, T b, acc c) {
    c.p = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c.p, 0, 0, 0);
    return c;
  }
  static __device__ int row(int lane, int j) { return 4 * (lane >> 4) + j; }
  static __device__ void flush(acc &c) {
#pragma unroll
    for (int j = 0; j < 4; ++j) c.s[j] += double(c.p[j]);
    c.p = f4{0.f, 0.f, 0.f, 0.f};
  }
  static __device__ double val(const acc &c, int j) { return c.s[j] + double(c.p[j]); }
};

// v_mfma_f32_16x16x32_bf16: lane l supplies A[l%16][8*(l/16)+0..7] and B[8*(l/16)+0..7][l%16];
// D[4*(l/16) + j][l%16].
__device__ inline f4 mma_bf16(bf8 a, bf8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ inline int bf16_row(int lane, int j) { return 4 * (lane >> 4) + j; }

__device__ inline void split_bf16(float x, bf16 &hi, bf16 &lo) {
  hi = (bf16)x;
  lo = (bf16)(x - (float)hi);
}

// Canonical slot of AO pair (l, s): Jb*64 + local, with the larger index as the "r" of the block pair.
__device__ inline size_t slot_of(int l, int s) {
  const int hi = l > s ? l : s, lo = l > s ? s : l;
  return size_t(zq_bpair(hi >> 3, lo >> 3)) * 64 + (hi & 7) * 8 + (lo & 7);
}

// Where the 64 x 64 tile (mu in block bmu, nu in block bnu | ket pairs of block pair Jb) is stored.
struct Q1Src {
  size_t base;
  int Ib, Jb;
  bool swp, diagB, diagK;
};

__device__ inline Q1Src q1_src(int bmu, int bnu, int br, int bs, int Jb) {
  const int bmax = bmu > bnu ? bmu : bnu, bmin = bmu > bnu ? bnu : bmu;
  Q1Src s;
  s.Ib = zq_bpair(bmax, bmin);
  s.Jb = Jb;
  s.swp = bmu < bnu;
  s.diagB = bmax == bmin;
  s.diagK = br == bs;
  s.base = (s.Ib >= Jb ? zq_blk_index(bmax, bmin, br, bs) : zq_blk_index(br, bs, bmax, bmin)) * size_t(ZQ_QUARTET);
  return s;
}

// (mu = bmu*8+u, nu = bnu*8+w | K): 0 for non-canonical ket slots of a diagonal ket block.  E is the
// element type of the resident quartets (the FP64 set, or its FP32 copy for the FP32 forms).
template <class E>
__device__ inline E q1_val(const E *__restrict__ eri, const Q1Src &s, int K, int u, int w) {
  if (s.diagK && (K >> 3) < (K & 7)) return E(0);
  int pl = s.swp ? w : u, ql = s.swp ? u : w;
  if (s.diagB && pl < ql) {
    const int t = pl;
    pl = ql, ql = t;
  }
  const int bl = pl * 8 + ql;
  int idx;
  if (s.Ib > s.Jb)
    idx = K * 64 + bl;
  else if (s.Ib < s.Jb)
    idx = bl * 64 + K;
  else
    idx = (K < bl ? K : bl) * 64 + (K < bl ? bl : K);
  return eri[s.base + idx];
}

// Element index (within its quartet) of (mu = bmu*8+u, nu = bnu*8+w | K), branch-free.  Unlike q1_val the
// non-canonical ket slots of a diagonal ket block are not zeroed: Q1 zeroes their outputs when it stores X.
__device__ inline int q1_idx(const Q1Src &s, int K, int u, int w) {
  int pl = s.swp ? w : u, ql = s.swp ? u : w;
  const int hi = max(pl, ql), lo = min(pl, ql);
  if (s.diagB) pl = hi, ql = lo;
  const int bl = pl * 8 + ql;
  const int kbl = s.Ib > s.Jb ? K * 64 + bl : bl * 64 + K;
  return s.Ib == s.Jb ? min(K, bl) * 64 + max(K, bl) : kbl;
}

// ---------------------------------------------------------------- Q1, FP64 / FP32
// Workgroup (256 threads) owns half kh of ket block pair Jb's 64 slots (K = 32 kh .. 32 kh + 31) x AO
// block bnu (8 nu) x a group of 16 MT occupied orbitals.  Loop over mu blocks: stage
// T[K][u*8+w] = (mu=bmu*8+u, nu=bnu*8+w | K) into LDS, then X[K][i][nu] += C1[mu,i] T via MFMA.  Wave w
// owns the 16-column tiles 4w .. 4w+3 (K = 32 kh + 8w .. 8w+7, all 8 nu) for all MT row tiles, so each
// LDS operand feeds MT MFMAs.
//
// On gfx950 the FP64 MFMAs do not co-issue with other vector ALU instructions, so the loop is written to
// issue few of them:
// - a thread's eight staging elements of a quartet block lie at one per-thread offset plus a uniform
//   stride, for each storage order (bra-fastest [K][bra] or ket-fastest [bra][K]) and bra orientation;
//   the offsets are computed once, and only the diagonal forms (Ib == Jb, or a diagonal bra block stored
//   ket-fastest) compute indices element by element;
// - LDS stores and operand reads use immediate offsets from fixed per-thread bases;
// - the non-canonical ket slots of a diagonal ket block are zeroed in the epilogue (X there is exactly
//   +0, as the zeroed inputs gave).
// Quartet blocks are fetched two mu blocks ahead and staged one ahead (double-buffered LDS, one barrier
// per block); the C1 operands of a block are loaded one block ahead, before the quartet fetch.  Loads,
// fetches and stages are unconditional (indices past the last block are clamped: their copies land in
// the buffer no later step reads, and the A-operand rows past m1 load row m1 - 1, which only reaches output
// rows that are never stored), so every path through the loop leaves the same loads in flight and the
// compiler's memory-counter waits are exact: the MFMAs wait only for their C1 operands, a stage only for
// its own quartet block.
// The body takes the workgroup's coordinates (bx, by, bz) as arguments, so that a grouped launch over
// many molecules can run it for each molecule's share of the grid.
struct Q1Slots {
  int kA, uA, wA, kB, wB;  // bra-fastest slot (kA + 4j, uA, wA); ket-fastest slot (kB, j, wB), j = 0..7
  int offA0, offA1, offAd, offB0, offB1;
};

// Loads this thread's eight elements of block (bm, bnu | Jb); returns true for ket-fastest slots.  Only the
// element indices depend on the block's form, the loads are one sequence (separate load paths would make
// the compiler merge their memory-counter states conservatively).  The slots are taken by value: a select
// between the addresses of their fields would keep them in scratch memory.
template <class T, class E>
__device__ __forceinline__ bool q1_fetch(const E *__restrict__ eri, const Q1Src src, const Q1Slots sl, int K0,
                                         T (&v)[8]) {
  const E *q = eri + src.base;
  const bool bra = src.Ib > src.Jb, ket = src.Ib < src.Jb;
  int idx[8];
  if (bra || (ket && !src.diagB)) {
    const int offA = sl.offA0 + (src.swp ? sl.offA1 - sl.offA0 : 0) + (src.diagB ? sl.offAd - sl.offA0 : 0);
    const int offK = src.swp ? sl.offB1 + 0 : sl.offB0 + 0;
    const int off = bra ? offA : offK;
    const int stride = bra ? 256 : (src.swp ? 64 : 512);
#pragma unroll
    for (int j = 0; j < 8; ++j) idx[j] = off + stride * j;
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int kl = ket ? sl.kB + 0 : sl.kA + 4 * j, u = ket ? j : sl.uA + 0, w = ket ? sl.wB + 0 : sl.wA + 0;
      idx[j] = q1_idx(src, K0 + kl, u, w);
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = T(q[idx[j]]);
  return ket;
}

// T[kl][u*8+w] of the staged block: bra-fastest slots at lA + 4 rows per j, ket-fastest at lB + 8 per j.
template <class T>
__device__ __forceinline__ void q1_stage(T *lA, T *lB, const T (&v)[8], bool ket) {
  if (ket) {
#pragma unroll
    for (int j = 0; j < 8; ++j) lB[8 * j] = v[j];
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) lA[4 * 65 * j] = v[j];
  }
}

// A operands of one mu block: C1[8 bm + 4 ks + (lane >> 4)][i0 + 16 h + (lane & 15)], cb at the block's row
// and coff[h] the (clamped) column of row tile h.
template <class T, int MT>
__device__ __forceinline__ void q1_load_c(const T *__restrict__ cb, size_t ldc1, const int (&coff)[MT], T (&cv)[2][MT]) {
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int h = 0; h < MT; ++h) cv[ks][h] = cb[ks * 4 * ldc1 + coff[h]];
}

// The block's MFMAs: B operand of column tile c is T[8w + 2c + ((lane >> 3) & 1)][(4 ks + (lane >> 4)) * 8
// + (lane & 7)], read from lR + 130 c + 32 ks.
template <class MF, int MT>
__device__ __forceinline__ void q1_mfmas(const typename MF::T *lR, const typename MF::T (&cv)[2][MT],
                                         typename MF::acc (&acc)[MT][4]) {
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const typename MF::T b = lR[2 * c * 65 + 32 * ks];
#pragma unroll
      for (int h = 0; h < MT; ++h) acc[h][c] = MF::mma(cv[ks][h], b, acc[h][c]);
    }
#pragma unroll
  for (int h = 0; h < MT; ++h)
#pragma unroll
    for (int c = 0; c < 4; ++c) MF::flush(acc[h][c]);
}

template <class MF, int MT, class E>
__device__ __forceinline__ void q1_body(const E *__restrict__ eri, const typename MF::T *__restrict__ C1, int ldc1,
                                        int m1, int nb, int npad, typename MF::T *__restrict__ X, int bx, int by,
                                        int bz) {
  typedef typename MF::T T;
  __shared__ T Ts[2][32][65];
  const int Jb = int(unsigned(bx) / unsigned(nb)), bnu = int(unsigned(bx) % unsigned(nb));
  const int i0 = by * (MT * 16);
  const int K0 = 32 * bz;
  int br, bs;
  zq_unpair(Jb, br, bs);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  typename MF::acc acc[MT][4];
#pragma unroll
  for (int h = 0; h < MT; ++h)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[h][c] = MF::zero();
  // staging slots: bra-fastest (K0 + kA + 4j, uA, wA) at (K0+kl)*64 + pl*8 + ql, (pl, ql) = (u, w), (w, u)
  // when the bra block is swapped, (max, min) on a diagonal bra block; ket-fastest (K0 + kB, j, wB) at
  // (pl*8 + ql)*64 + K
  Q1Slots sl;
  sl.kA = tid >> 6, sl.uA = (tid >> 3) & 7, sl.wA = tid & 7;
  sl.kB = tid & 31, sl.wB = tid >> 5;
  sl.offA0 = (K0 + sl.kA) * 64 + sl.uA * 8 + sl.wA;
  sl.offA1 = (K0 + sl.kA) * 64 + sl.wA * 8 + sl.uA;
  sl.offAd = (K0 + sl.kA) * 64 + max(sl.uA, sl.wA) * 8 + min(sl.uA, sl.wA);
  sl.offB0 = sl.wB * 64 + K0 + sl.kB;
  sl.offB1 = sl.wB * 512 + K0 + sl.kB;
  T *const lA0 = &Ts[0][sl.kA][tid & 63], *const lA1 = &Ts[1][sl.kA][tid & 63];
  T *const lB0 = &Ts[0][sl.kB][sl.wB], *const lB1 = &Ts[1][sl.kB][sl.wB];
  const T *const lR0 = &Ts[0][8 * wave + ((lane >> 3) & 1)][(lane >> 4) * 8 + (lane & 7)];
  const T *const lR1 = &Ts[1][8 * wave + ((lane >> 3) & 1)][(lane >> 4) * 8 + (lane & 7)];
  const T *const cbase = C1 + size_t(lane >> 4) * ldc1;
  const size_t ldc = size_t(ldc1);
  int coff[MT];
#pragma unroll
  for (int h = 0; h < MT; ++h) coff[h] = min(i0 + 16 * h + (lane & 15), m1 - 1);

  T ca[2][MT], cb[2][MT], ra[8], rb[8];
  const int last = nb - 1;
  q1_load_c<T, MT>(cbase, ldc, coff, ca);
  q1_stage(lA0, lB0, ra, q1_fetch<T, E>(eri, q1_src(0, bnu, br, bs, Jb), sl, K0, ra));
  bool kb = q1_fetch<T, E>(eri, q1_src(min(1, last), bnu, br, bs, Jb), sl, K0, rb);
  __syncthreads();
  for (int bm = 0; bm < nb; bm += 2) {
    // block bm from buffer 0 with ca; fetch bm + 2 into ra; stage bm + 1 (rb) into buffer 1
    q1_load_c<T, MT>(cbase + size_t(min(bm + 1, last)) * 8 * ldc, ldc, coff, cb);
    const bool ka = q1_fetch<T, E>(eri, q1_src(min(bm + 2, last), bnu, br, bs, Jb), sl, K0, ra);
    q1_mfmas<MF, MT>(lR0, ca, acc);
    q1_stage(lA1, lB1, rb, kb);
    __syncthreads();
    // block bm + 1 from buffer 1 with cb; fetch bm + 3 into rb; stage bm + 2 (ra) into buffer 0
    q1_load_c<T, MT>(cbase + size_t(min(bm + 2, last)) * 8 * ldc, ldc, coff, ca);
    kb = q1_fetch<T, E>(eri, q1_src(min(bm + 3, last), bnu, br, bs, Jb), sl, K0, rb);
    if (bm + 1 < nb) q1_mfmas<MF, MT>(lR1, cb, acc);
    q1_stage(lA0, lB0, ra, ka);
    __syncthreads();
  }
  const bool diagK = br == bs;
#pragma unroll
  for (int h = 0; h < MT; ++h)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int col = (4 * wave + c) * 16 + (lane & 15), K = K0 + (col >> 3), wl = col & 7;
      const bool zero = diagK && (K >> 3) < (K & 7);  // non-canonical slot of a diagonal ket block
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = i0 + h * 16 + MF::row(lane, j);
        if (i < m1)
          X[(size_t(Jb) * 64 + K) * size_t(m1) * npad + size_t(i) * npad + bnu * 8 + wl] = zero ? T(0) : T(MF::val(acc[h][c], j));
      }
    }
}

// Two workgroups per CU (the accumulators then sit in the architectural VGPRs), unless the accumulators
// alone take more than 144 of a wave's registers (FP32 products with FP64 sums at MT = 4), which would spill.
template <class MF, int MT>
constexpr int q1_min_waves() {
  return MT * 4 * MF::kAccRegs <= 144 ? 2 : 1;
}

template <class MF, int MT, class E>
__global__ __launch_bounds__(256, (q1_min_waves<MF, MT>())) void zq_q1(const E *__restrict__ eri, const typename MF::T *__restrict__ C1,
                                                int ldc1, int m1, int nb, int npad, typename MF::T *__restrict__ X) {
  q1_body<MF, MT, E>(eri, C1, ldc1, m1, nb, npad, X, blockIdx.x, blockIdx.y, blockIdx.z);
}

// ---------------------------------------------------------------- Q1, BF16X3
// Same tiling, four mu blocks (32 mu, one bf16 MFMA k-step) staged per step as hi/lo halves
// Th/Tl[K][w][mu_local].  C1 comes pre-split and transposed: C1h/C1l[i][mu] (ldt >= 8*ceil(nb/4)*4).
template <int MT>
__global__ __launch_bounds__(512) void zq_q1_bf16x3(const double *__restrict__ eri, const bf16 *__restrict__ C1h,
                                                    const bf16 *__restrict__ C1l, int ldt, int m1, int nb, int npad,
                                                    float *__restrict__ X) {
  constexpr int TPW = MT * 4;
  __shared__ __attribute__((aligned(16))) bf16 Th[64][8][40];
  __shared__ __attribute__((aligned(16))) bf16 Tl[64][8][40];
  const int Jb = blockIdx.x / nb, bnu = blockIdx.x % nb;
  const int i0 = blockIdx.y * 64;
  int br, bs;
  zq_unpair(Jb, br, bs);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  f4 big[TPW], small[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) big[t] = small[t] = f4{0.f, 0.f, 0.f, 0.f};

  for (int bm0 = 0; bm0 < nb; bm0 += 4) {
    __syncthreads();
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
      const int bmu = bm0 + bi;
      if (bmu < nb) {
        const Q1Src src = q1_src(bmu, bnu, br, bs, Jb);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int e = tid + 512 * j;
          const int K = e >> 6, uw = e & 63, u = uw >> 3, w = uw & 7;
          bf16 h, l;
          split_bf16(float(q1_val<double>(eri, src, K, u, w)), h, l);
          Th[K][w][bi * 8 + u] = h;
          Tl[K][w][bi * 8 + u] = l;
        }
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int e = tid + 512 * j;
          const int K = e >> 6, uw = e & 63;
          Th[K][uw & 7][bi * 8 + (uw >> 3)] = (bf16)0.f;
          Tl[K][uw & 7][bi * 8 + (uw >> 3)] = (bf16)0.f;
        }
      }
    }
    __syncthreads();
    const int kg = lane >> 4;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int tile = wave * TPW + t, it = tile >> 5, nt = tile & 31;
      const int i = i0 + it * 16 + (lane & 15);
      bf8 ah, al;
      if (i < m1) {
        ah = *(const bf8 *)(C1h + size_t(i) * ldt + bm0 * 8 + kg * 8);
        al = *(const bf8 *)(C1l + size_t(i) * ldt + bm0 * 8 + kg * 8);
      } else {
#pragma unroll
        for (int x = 0; x < 8; ++x) ah[x] = al[x] = (bf16)0.f;
      }
      const int col = nt * 16 + (lane & 15), K = col >> 3, wl = col & 7;
      const bf8 bh = *(const bf8 *)&Th[K][wl][kg * 8];
      const bf8 bl = *(const bf8 *)&Tl[K][wl][kg * 8];
      big[t] = mma_bf16(ah, bh, big[t]);
      small[t] = mma_bf16(ah, bl, small[t]);
      small[t] = mma_bf16(al, bh, small[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = wave * TPW + t, it = tile >> 5, nt = tile & 31;
    const int col = nt * 16 + (lane & 15), K = col >> 3, wl = col & 7;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = i0 + it * 16 + bf16_row(lane, j);
      if (i < m1) X[(size_t(Jb) * 64 + K) * size_t(m1) * npad + size_t(i) * npad + bnu * 8 + wl] = big[t][j] + small[t][j];
    }
  }
}

// ---------------------------------------------------------------- batched GEMM
// C[b][m][n] = sum_k A[b][m][k] B[b][k][n] with arbitrary element strides.  Mode 1 gathers B from X:
// B[lam][col = sig*npad + nu] = X[slot(lam, sig)][b][nu] (slot stride sXs; B already offset by b).
// Batch entry b = b0 + blockIdx.z.  tri 0: A, B, C at b sAb, b sBb, b sCb.  tri 1: the same, with rows
// m <= b only (Q2 over j <= i).  tri 2: b is the pair index p = i (i + 1) / 2 + j of i >= j, A at
// i sAi + j sAj (Q3 over the pairs of a symmetric transform).
template <class T>
struct GemmArgs {
  const T *A;
  long long sAb, sAm, sAk;
  const T *B;
  long long sBb, sBk, sBn;
  T *C;
  long long sCb, sCm, sCn;
  int M, N, K;
  int mode;
  int npad;
  long long sXs;
  int tri, b0;
  long long sAi, sAj;
};

// (i, j), i >= j, of the pair index p = i (i + 1) / 2 + j.
__device__ inline void pair_of(int p, int &i, int &j) {
  int r = int((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
  while ((r + 1) * (r + 2) / 2 <= p) ++r;
  while (r * (r + 1) / 2 > p) --r;
  i = r, j = p - r * (r + 1) / 2;
}

// Operand bases and the row count of batch entry b.
template <class T>
__device__ inline void gemm_batch(const GemmArgs<T> &g, int b, const T *&A, const T *&B, T *&C, int &M) {
  M = g.M;
  if (g.tri == 2) {
    int i, j;
    pair_of(b, i, j);
    A = g.A + i * g.sAi + j * g.sAj;
  } else {
    A = g.A + b * g.sAb;
    if (g.tri == 1) M = min(g.M, b + 1);
  }
  B = g.B + b * g.sBb;
  C = g.C + b * g.sCb;
}

template <class T>
__device__ inline T gemm_b(const GemmArgs<T> &g, const T *B, int k, int n) {
  if (g.mode == 1) {
    const int sig = n / g.npad, nu = n - sig * g.npad;
    return B[(long long)slot_of(k, sig) * g.sXs + nu];
  }
  return B[k * g.sBk + n * g.sBn];
}

// FP64 / FP32 / FP32C: 64 x 64 tiles, BK = 16, four waves of 32 x 32 (2 x 2 MFMA tiles, so each LDS
// operand read feeds two MFMAs).  Written, like Q1, to issue few vector ALU instructions (the FP64 MFMAs
// do not co-issue with them):
// - each thread loads 4 elements of A and 4 of B per K step at per-thread offsets fixed for the whole
//   loop, from uniform K-step bases; rows past M and columns past N are clamped to the last one (they
//   only reach outputs that are not stored), so no load is masked;
// - the gathered B of Q2 (mode 1, element (lambda, (sigma, nu)) at X[slot(lambda, sigma)][i][nu]) picks,
//   per element, one of three per-thread offsets by whether lambda's AO block lies above, below or at
//   sigma's, plus a uniform part per lambda block;
// - K steps are loaded two ahead and staged one ahead, unconditionally (step indices past the last are
//   clamped and their copies land in the LDS buffer no later step reads), so every path leaves the same
//   loads in flight and the compiler's memory-counter waits are exact;
// - a K that is not a multiple of 16 ends with one masked step (k >= K loads k = K - 1 and stages 0).
struct GemmSlots {
  unsigned offA[4], offB[4];           // mode 0: element offsets from the K-step bases (within one matrix)
  unsigned long long p1[4], p2[4], pe[4];  // mode 1: slot offsets, lambda's block above / below / at sigma's
  int kA[4], kB[4];                    // k of element j within its step
  int mA[4], nB[4];                    // row of A element j, column of B element j (clamped)
};

template <class MF>
__device__ __forceinline__ void gemm_load(const GemmArgs<typename MF::T> &g, const typename MF::T *A,
                                          const typename MF::T *B, const GemmSlots sl, int k0, int sb,
                                          typename MF::T (&ra)[4], typename MF::T (&rb)[4]) {
  const typename MF::T *Ak = A + (long long)k0 * g.sAk;
#pragma unroll
  for (int j = 0; j < 4; ++j) ra[j] = Ak[sl.offA[j]];
  if (g.mode == 1) {
    const unsigned long long sX = (unsigned long long)g.sXs;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int lb = (k0 >> 3) + h;  // lambda block of elements 2h, 2h + 1 (uniform)
      const unsigned long long u1 = (unsigned long long)(lb * (lb + 1) / 2) * 64 * sX;  // block pair (lb, sb)
      const unsigned long long u2 = (unsigned long long)lb * 64 * sX;                    // block pair (sb, lb)
      const int rel = lb - sb;
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int j = 2 * h + q;
        const unsigned long long off = rel > 0 ? u1 + sl.p1[j] : (rel < 0 ? u2 + sl.p2[j] : sl.pe[j]);
        rb[j] = B[off];
      }
    }
  } else {
    const typename MF::T *Bk = B + (long long)k0 * g.sBk;
#pragma unroll
    for (int j = 0; j < 4; ++j) rb[j] = Bk[sl.offB[j]];
  }
}

// The last, partial K step: elements with k >= K are zero (their loads read k = K - 1).
template <class MF>
__device__ __forceinline__ void gemm_load_tail(const GemmArgs<typename MF::T> &g, const typename MF::T *A,
                                               const typename MF::T *B, const GemmSlots sl, int k0,
                                               typename MF::T (&ra)[4], typename MF::T (&rb)[4]) {
  typedef typename MF::T T;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = k0 + sl.kA[j], kc = min(k, g.K - 1);
    const T x = A[(long long)sl.mA[j] * g.sAm + (long long)kc * g.sAk];
    ra[j] = k < g.K ? x : T(0);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = k0 + sl.kB[j], kc = min(k, g.K - 1);
    const T x = gemm_b(g, B, kc, sl.nB[j]);
    rb[j] = k < g.K ? x : T(0);
  }
}

// a_kfast: A element j at As[k][m + 16 j], else As[k + 4 j][m]; b_nfast: B element j at Bs[k + 4 j][n],
// else Bs[k][n + 16 j] (lA, lB: the thread's element 0 in the buffer)
template <class T>
__device__ __forceinline__ void gemm_stage(T *lA, T *lB, bool a_kfast, bool b_nfast, const T (&ra)[4],
                                           const T (&rb)[4]) {
  if (a_kfast) {
#pragma unroll
    for (int j = 0; j < 4; ++j) lA[16 * j] = ra[j];
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) lA[4 * 65 * j] = ra[j];
  }
  if (b_nfast) {
#pragma unroll
    for (int j = 0; j < 4; ++j) lB[4 * 65 * j] = rb[j];
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) lB[16 * j] = rb[j];
  }
}

// One K step: operands As[4 ks + (lane >> 4)][wm + 16 x + (lane & 15)] and the same of Bs (lRA, lRB at
// ks = x = 0).  16 x 16 sub-tiles wholly outside M or N take no MFMA (uniform over the wave).
template <class MF>
__device__ __forceinline__ void gemm_mfmas(const typename MF::T *lRA, const typename MF::T *lRB,
                                           const bool (&live)[2][2], typename MF::acc (&acc)[2][2]) {
  typedef typename MF::T T;
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    T a[2], b[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) a[x] = lRA[ks * 4 * 65 + 16 * x];
#pragma unroll
    for (int y = 0; y < 2; ++y) b[y] = lRB[ks * 4 * 65 + 16 * y];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y)
        if (live[x][y]) acc[x][y] = MF::mma(a[x], b[y], acc[x][y]);
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) MF::flush(acc[x][y]);
}

// Workgroup (bx, by, bz): the 64 x 64 output tile (by, bx) of batch entry bz.
template <class MF>
__device__ __forceinline__ void gemm_body(const GemmArgs<typename MF::T> &g, int bx, int by, int bz) {
  typedef typename MF::T T;
  __shared__ T As[2][16][65];
  __shared__ T Bs[2][16][65];
  const int m0 = by * 64, n0 = bx * 64;
  const T *A, *B;
  T *C;
  int M;
  gemm_batch(g, g.b0 + bz, A, B, C, M);
  if (m0 >= M) return;  // whole workgroup: rows of a tri 1 entry beyond its pair
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  typename MF::acc acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = MF::zero();
  const bool a_kfast = (g.sAk == 1);
  const bool b_nfast = (g.mode == 1) || (g.sBn == 1);
  bool live[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) live[x][y] = m0 + wm + 16 * x < M && n0 + wn + 16 * y < g.N;

  // the thread's element j: A (m, k) = (m0 + (tid >> 4) + 16 j, tid & 15) if a_kfast, else (m0 + (tid & 63),
  // (tid >> 6) + 4 j); B (k, n) = ((tid >> 6) + 4 j, n0 + (tid & 63)) if b_nfast, else (tid & 15,
  // n0 + (tid >> 4) + 16 j)
  GemmSlots sl;
  int sb = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int mm = a_kfast ? (tid >> 4) + 16 * j : tid & 63, kk = a_kfast ? tid & 15 : (tid >> 6) + 4 * j;
    const int m = min(m0 + mm, M - 1);
    sl.mA[j] = m, sl.kA[j] = kk;
    sl.offA[j] = unsigned((long long)m * g.sAm + (long long)kk * g.sAk);
    const int nn = b_nfast ? tid & 63 : (tid >> 4) + 16 * j, kb = b_nfast ? (tid >> 6) + 4 * j : tid & 15;
    const int n = min(n0 + nn, g.N - 1);
    sl.nB[j] = n, sl.kB[j] = kb;
    sl.offB[j] = unsigned((long long)kb * g.sBk + (long long)n * g.sBn);
  }
  if (g.mode == 1) {
    // slot(lambda, sigma) = bpair(max block, min block) * 64 + (max & 7) * 8 + (min & 7), as slot_of
    const int n = sl.nB[0], sig = n / g.npad, nu = n - sig * g.npad, sr = sig & 7;
    sb = sig >> 3;
    const unsigned long long sX = (unsigned long long)g.sXs;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int lr = ((tid >> 6) + 4 * j) & 7;
      sl.p1[j] = (unsigned long long)(sb * 64 + lr * 8 + sr) * sX + nu;
      sl.p2[j] = (unsigned long long)((sb * (sb + 1) / 2) * 64 + sr * 8 + lr) * sX + nu;
      sl.pe[j] = (unsigned long long)((sb * (sb + 1) / 2 + sb) * 64 + max(lr, sr) * 8 + min(lr, sr)) * sX + nu;
    }
  }
  T *const lA0 = a_kfast ? &As[0][tid & 15][tid >> 4] : &As[0][tid >> 6][tid & 63], *const lA1 = lA0 + 16 * 65;
  T *const lB0 = b_nfast ? &Bs[0][tid >> 6][tid & 63] : &Bs[0][tid & 15][tid >> 4], *const lB1 = lB0 + 16 * 65;
  const T *const lRA0 = &As[0][lane >> 4][wm + (lane & 15)], *const lRA1 = lRA0 + 16 * 65;
  const T *const lRB0 = &Bs[0][lane >> 4][wn + (lane & 15)], *const lRB1 = lRB0 + 16 * 65;

  const int nfull = g.K >> 4, last = nfull - 1;
  T ra[4], qa[4], rb[4], qb[4];  // (A, B) of the step in flight by parity: (ra, qa) even, (rb, qb) odd
  if (nfull > 0) {
    gemm_load<MF>(g, A, B, sl, 0, sb, ra, qa);
    gemm_stage(lA0, lB0, a_kfast, b_nfast, ra, qa);
    gemm_load<MF>(g, A, B, sl, 16 * min(1, last), sb, rb, qb);
    __syncthreads();
    for (int s = 0; s < nfull; s += 2) {
      // step s from buffer 0; load s + 2 into (ra, qa); stage s + 1 (rb, qb) into buffer 1
      gemm_load<MF>(g, A, B, sl, 16 * min(s + 2, last), sb, ra, qa);
      gemm_mfmas<MF>(lRA0, lRB0, live, acc);
      gemm_stage(lA1, lB1, a_kfast, b_nfast, rb, qb);
      __syncthreads();
      // step s + 1 from buffer 1; load s + 3 into (rb, qb); stage s + 2 (ra, qa) into buffer 0
      gemm_load<MF>(g, A, B, sl, 16 * min(s + 3, last), sb, rb, qb);
      if (s + 1 < nfull) gemm_mfmas<MF>(lRA1, lRB1, live, acc);
      gemm_stage(lA0, lB0, a_kfast, b_nfast, ra, qa);
      __syncthreads();
    }
  }
  if (g.K & 15) {  // the partial last step, through buffer 0 (no step reads it any more)
    gemm_load_tail<MF>(g, A, B, sl, 16 * nfull, ra, qa);
    gemm_stage(lA0, lB0, a_kfast, b_nfast, ra, qa);
    __syncthreads();
    gemm_mfmas<MF>(lRA0, lRB0, live, acc);
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int n = n0 + wn + y * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int m = m0 + wm + x * 16 + MF::row(lane, j);
        if (m < M && n < g.N) C[m * g.sCm + n * g.sCn] = T(MF::val(acc[x][y], j));
      }
    }
}

template <class MF>
__global__ __launch_bounds__(256, 2) void zq_gemm(GemmArgs<typename MF::T> g) {
  gemm_body<MF>(g, blockIdx.x, blockIdx.y, blockIdx.z);
}

// BF16X3: FP32 operands split into bf16 hi/lo while staged; 64 x 64 tiles, BK = 32 (one k-step).
__global__ __launch_bounds__(256) void zq_gemm_bf16x3(GemmArgs<float> g) {
  __shared__ __attribute__((aligned(16))) bf16 Ah[64][40];
  __shared__ __attribute__((aligned(16))) bf16 Al[64][40];
  __shared__ __attribute__((aligned(16))) bf16 Bh[64][40];
  __shared__ __attribute__((aligned(16))) bf16 Bl[64][40];
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const float *A, *B;
  float *C;
  int M;
  gemm_batch(g, g.b0 + int(blockIdx.z), A, B, C, M);
  if (m0 >= M) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f4 big[2][2], small[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) big[x][y] = small[x][y] = f4{0.f, 0.f, 0.f, 0.f};
  const bool a_kfast = (g.sAk == 1);
  const bool b_nfast = (g.mode == 1) || (g.sBn == 1);
  for (int k0 = 0; k0 < g.K; k0 += 32) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = tid + 256 * j;
      int mm, kk;
      if (a_kfast)
        mm = e >> 5, kk = e & 31;
      else
        kk = e >> 6, mm = e & 63;
      const int m = m0 + mm, k = k0 + kk;
      const float v = (m < g.M && k < g.K) ? A[m * g.sAm + k * g.sAk] : 0.f;
      split_bf16(v, Ah[mm][kk], Al[mm][kk]);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = tid + 256 * j;
      int nn, kk;
      if (b_nfast)
        kk = e >> 6, nn = e & 63;
      else
        nn = e >> 5, kk = e & 31;
      const int n = n0 + nn, k = k0 + kk;
      const float v = (n < g.N && k < g.K) ? gemm_b(g, B, k, n) : 0.f;
      split_bf16(v, Bh[nn][kk], Bl[nn][kk]);
    }
    __syncthreads();
    const int kg = (lane >> 4) * 8;
    bf8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      ah[x] = *(const bf8 *)&Ah[wm + x * 16 + (lane & 15)][kg];
      al[x] = *(const bf8 *)&Al[wm + x * 16 + (lane & 15)][kg];
      bh[x] = *(const bf8 *)&Bh[wn + x * 16 + (lane & 15)][kg];
      bl[x] = *(const bf8 *)&Bl[wn + x * 16 + (lane & 15)][kg];
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        big[x][y] = mma_bf16(ah[x], bh[y], big[x][y]);
        small[x][y] = mma_bf16(ah[x], bl[y], small[x][y]);
        small[x][y] = mma_bf16(al[x], bh[y], small[x][y]);
      }
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int n = n0 + wn + y * 16 + (lane & 15);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int m = m0 + wm + x * 16 + bf16_row(lane, j);
        if (m < M && n < g.N) C[m * g.sCm + n * g.sCn] = big[x][y][j] + small[x][y][j];
      }
    }
}

// ---------------------------------------------------------------- MP2 energy (FP64 arithmetic)
// One workgroup per (i,j): sum_ab T[a][b] (2 T[a][b] - T[b][a]) / (e_i + e_j - e_a - e_b).  The v x v
// block is taken in 32 x 32 tile pairs (A <= B): tiles T[A][B] and T[B][A] are staged through LDS along
// their rows, and each (a,b), a < b, contributes with its mirror 2 (x^2 - x y + y^2) / d (x = T[a][b],
// y = T[b][a]), a diagonal element x^2 / d.  Every thread accumulates its positions in a fixed order and
// the threads are summed in a fixed tree: partial[ij], bitwise reproducible.
// packed: ij is the pair index of i >= j (blocks of the symmetric transform's pairs only) and an i > j
// block counts twice, for its mirror (j,i), whose sum is the same (its block is this one transposed and
// its denominators are those of (b,a)); otherwise ij = i o + j.
template <class T>
__device__ __forceinline__ void mp2_pairs_body(const T *__restrict__ t4, const double *__restrict__ e, int o, int v,
                                               double *__restrict__ partial, int ij, bool packed) {
  __shared__ double X[32][33], Yt[32][33];
  __shared__ double red[256];
  int i, j;
  if (packed)
    pair_of(ij, i, j);
  else
    i = ij / o, j = ij % o;
  const T *t = t4 + size_t(ij) * v * v;
  const double eij = e[i] + e[j];
  const double *ev = e + o;
  const int tid = threadIdx.x, c = tid & 31, r0 = tid >> 5;  // r0 = 0..7
  const int nt = (v + 31) / 32;
  double s = 0.0;
  for (int A = 0; A < nt; ++A)
    for (int B = A; B < nt; ++B) {
      __syncthreads();  // the previous pair's tiles are consumed
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int r = r0 + 8 * k;
        const int a = A * 32 + r, b = B * 32 + c, b2 = B * 32 + r, a2 = A * 32 + c;
        X[r][c] = (a < v && b < v) ? double(t[size_t(a) * v + b]) : 0.0;    // T[A*32 + r][B*32 + c]
        Yt[r][c] = (b2 < v && a2 < v) ? double(t[size_t(b2) * v + a2]) : 0.0;  // T[B*32 + r][A*32 + c]
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int r = r0 + 8 * k;
        const int a = A * 32 + r, b = B * 32 + c;
        if (a >= v || b >= v) continue;
        const double x = X[r][c], y = Yt[c][r];  // T[a][b], T[b][a]
        const double d = eij - ev[a] - ev[b];
        if (A < B || r < c)
          s += 2.0 * (x * x - x * y + y * y) / d;
        else if (r == c)
          s += x * x / d;
      }
    }
  red[tid] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) partial[ij] = (packed && i != j) ? 2.0 * red[0] : red[0];
}

template <class T>
__global__ __launch_bounds__(256) void zq_mp2_pairs(const T *__restrict__ t4, const double *__restrict__ e, int o, int v,
                                                    double *__restrict__ partial, int packed) {
  mp2_pairs_body<T>(t4, e, o, v, partial, blockIdx.x, packed != 0);
}

// The full out[i][j][a][b] (m1 x m1 x m3 x m3) of a symmetric transform from its pair blocks O[p][a][b],
// i >= j: out[i][j] = O[p(i,j)], out[j][i] = O[p(i,j)] transposed.  One workgroup per (i,j).
template <class T>
__global__ __launch_bounds__(256) void zq_expand_pairs(const T *__restrict__ O, int m1, int m3, T *__restrict__ out) {
  const int ij = blockIdx.x, i = ij / m1, j = ij - i * m1;
  const bool lower = i >= j;
  const int p = lower ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
  const T *src = O + size_t(p) * m3 * m3;
  T *dst = out + size_t(ij) * m3 * m3;
  const int mm = m3 * m3;
  for (int x = threadIdx.x; x < mm; x += 256) {
    const int a = x / m3, b = x - a * m3;
    dst[x] = lower ? src[x] : src[size_t(b) * m3 + a];
  }
}

// *out = sum of x[0..n): lane l adds x[l], x[l + 64], ... in order, then a fixed butterfly.
__device__ __forceinline__ void sum_wave_body(const double *__restrict__ x, int n, double *__restrict__ out) {
  double s = 0.0;
  for (int k = threadIdx.x; k < n; k += 64) s += x[k];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (threadIdx.x == 0) *out = s;
}

__global__ __launch_bounds__(64) void zq_sum_wave(const double *__restrict__ x, int n, double *__restrict__ out) {
  sum_wave_body(x, n, out);
}

__global__ void zq_widen(const float *__restrict__ x, size_t n, double *__restrict__ y) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) y[i] = x[i];
}

// The batch index of the GEMM kernels is blockIdx.z, at most kMaxGridZ per launch: larger batches (the
// FULL transform's n^2 (i,j) pairs from n = 256 on) run as several launches over consecutive slices.
template <class T>
GemmArgs<T> batch_slice(GemmArgs<T> g, long long b0) {
  g.b0 = int(b0);
  return g;
}

template <class MF>
int launch_gemm(zq_ctx *ctx, const GemmArgs<typename MF::T> &g, int batch, const char *name) {
  zq_prof_begin(ctx, name);
  for (int b0 = 0; b0 < batch; b0 += kMaxGridZ) {
    const dim3 grid(unsigned((g.N + 63) / 64), unsigned((g.M + 63) / 64), unsigned(std::min(kMaxGridZ, batch - b0)));
    hipLaunchKernelGGL((zq_gemm<MF>), grid, dim3(256), 0, ctx->stream, batch_slice(g, b0));
  }
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, name);
  return ZQ_OK;
}

int launch_gemm_bf16x3(zq_ctx *ctx, const GemmArgs<float> &g, int batch, const char *name) {
  zq_prof_begin(ctx, name);
  for (int b0 = 0; b0 < batch; b0 += kMaxGridZ) {
    const dim3 grid(unsigned((g.N + 63) / 64), unsigned((g.M + 63) / 64), unsigned(std::min(kMaxGridZ, batch - b0)));
    hipLaunchKernelGGL(zq_gemm_bf16x3, grid, dim3(256), 0, ctx->stream, batch_slice(g, b0));
  }
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, name);
  return ZQ_OK;
}

// Device buffers of one transform: the context's grow-only scratch, or (base != null) consecutive
// 256-byte aligned pieces of the static no-allocation scratch, cap bytes long.
struct Arena {
  zq_ctx *ctx;
  char *base = nullptr;
  size_t cap = 0, off = 0;
  void *get(const char *name, size_t bytes) {
    if (!base) return zq_scratch(ctx, name, bytes, nullptr, false);
    const size_t at = (off + 255) & ~size_t(255);
    if (at + bytes > cap) return nullptr;
    off = at + bytes;
    return base + at;
  }
};

template <class T>
int upload(zq_ctx *ctx, const char *name, const std::vector<T> &h, T **d) {
  *d = (T *)zq_scratch(ctx, name, h.size() * sizeof(T), nullptr, false);
  if (!*d) return zq_fail(ctx, ZQ_ERR_OOM, std::string("transform scratch ") + name);
  ZQ_TRY(ctx, hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return ZQ_OK;
}

// The coefficient slices of run_chain built on the device from C (n x n, already on the device): the
// same element conversions as the host staging, so the values are identical.
template <class T>
__device__ __forceinline__ void c_slices_body(const double *__restrict__ C, int n, int np, int m1, int m2, int m3,
                                              int off3, T *__restrict__ c1, T *__restrict__ c2t, T *__restrict__ c3,
                                              int mu) {
  for (int t = threadIdx.x; t < m1 + m2 + m3; t += blockDim.x) {
    if (t < m1) {
      c1[size_t(mu) * m1 + t] = mu < n ? T(C[size_t(mu) * n + t]) : T(0);
    } else if (t < m1 + m2) {
      const int j = t - m1;
      c2t[size_t(j) * np + mu] = mu < n ? T(C[size_t(mu) * n + j]) : T(0);
    } else {
      const int a = t - m1 - m2;
      c3[size_t(mu) * m3 + a] = mu < n ? T(C[size_t(mu) * n + off3 + a]) : T(0);
    }
  }
}

template <class T>
__global__ void zq_c_slices(const double *__restrict__ C, int n, int np, int m1, int m2, int m3, int off3,
                            T *__restrict__ c1, T *__restrict__ c2t, T *__restrict__ c3) {
  c_slices_body<T>(C, n, np, m1, m2, m3, off3, c1, c2t, c3, blockIdx.x);  // one workgroup per row mu < np
}

// BF16X3 Q1 operand: C1 split into bf16 hi + lo and transposed, [i][mu] with leading dimension ldt.
__global__ void zq_c1_split(const double *__restrict__ C, int n, int m1, int ldt, bf16 *__restrict__ hh,
                            bf16 *__restrict__ hl) {
  const int i = blockIdx.x;  // m1 rows
  for (int mu = threadIdx.x; mu < ldt; mu += blockDim.x) {
    bf16 h = (bf16)0.f, l = (bf16)0.f;
    if (mu < n) {
      const float x = float(C[size_t(mu) * n + i]);
      h = (bf16)x;
      l = (bf16)(x - (float)h);
    }
    hh[size_t(i) * ldt + mu] = h;
    hl[size_t(i) * ldt + mu] = l;
  }
}

struct Dims {
  int n, np, nb, Nb, m1, m2, m3, m4, off3;
};

// Q1..Q4 with element type T (double: FP64, float: FP32 or BF16X3).
// The coefficient matrix of each quarter (mu->i, lam->j, nu->a, sig->b); all the same C except in the
// terms of the dual transform.
struct Coefs {
  const double *q[4];
  bool dev = false;  // q[] are device pointers (one matrix: the context's resident SCF orbitals)
};

// Bytes run_chain carves from a static arena for these dimensions and one coefficient matrix (a
// symmetric transform: Q3 and Q4 over the pairs i >= j), each piece 256-byte aligned: the
// no-allocation entry point's size check.
size_t chain_bytes(const Dims &d, int precision) {
  const size_t t = precision == ZQ_PREC_FP64 ? 8 : 4, np = d.np, m1 = d.m1, m2 = d.m2, m3 = d.m3, m4 = d.m4;
  const size_t ldt = ((d.nb + 3) / 4) * 32, pairs = m1 * (m1 + 1) / 2;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  size_t b = al(size_t(d.n) * d.n * 8) + al(np * m1 * t) + al(m2 * np * t) + al(np * m3 * t);
  if (precision == ZQ_PREC_BF16X3) b += 2 * al(m1 * ldt * sizeof(bf16));
  b += al(size_t(d.Nb) * 64 * m1 * np * t) + al(m1 * m2 * np * np * t) + al(pairs * np * m3 * t) + al(pairs * m3 * m4 * t);
  return b;
}

// Q1..Q4.  With a static arena (ar.base set) nothing is allocated: C goes to the arena and is sliced
// there, and FP32 forms need the resident FP32 ERI copy already made.
// A symmetric transform (the same coefficients for i and j, and for a and b) has out[j][i][b][a] =
// out[i][j][a][b], (jb|ia) = (ia|jb): Q2 then forms the rows j <= i only, Q3 and Q4 run over the pairs
// i >= j, and *d_out holds the pair blocks O[p][a][b], p = i (i + 1) / 2 + j (*packed = true).
// Otherwise *d_out is out[i][j][a][b] for all i, j.
template <class T, class MF>
int run_chain(zq_ctx *ctx, const Coefs &cf, const Dims &d, int precision, T **d_out, Arena &ar, bool *packed) {
  const int n = d.n, np = d.np, nb = d.nb, Nb = d.Nb, m1 = d.m1, m2 = d.m2, m3 = d.m3, m4 = d.m4;
  const bool fixed = ar.base != nullptr;
  const double *C = cf.q[0];
  const bool own4 = cf.q[3] != cf.q[2];  // a fourth quarter with its own coefficients (dual transform)
  if ((fixed || cf.dev) && (own4 || cf.q[1] != C || cf.q[2] != C))
    return zq_fail(ctx, ZQ_ERR_INVALID, "one coefficient matrix only");
  // BF16X3 Q1 takes C1 split and transposed: C1h/C1l[i][mu], mu padded to a multiple of 32
  const int ldt = ((nb + 3) / 4) * 32;
  T *d1, *d2t, *d3, *d4 = nullptr;
  bf16 *d1h = nullptr, *d1l = nullptr;
  if constexpr (sizeof(T) == 4) {
    if (precision != ZQ_PREC_BF16X3) {
      if (fixed && !ctx->eri32)
        return zq_fail(ctx, ZQ_ERR_STATE, "zq_mp2_noalloc: FP32 forms need the FP32 ERI copy resident (an ERI set made "
                                          "with_fp32, or an earlier FP32 call on this context)");
      if (int rc = zq_eri32(ctx)) return rc;
    }
  }
  if (fixed) {
    auto *dC = (double *)ar.get("C", size_t(n) * n * 8);
    d1 = (T *)ar.get("C1", size_t(np) * m1 * sizeof(T));
    d2t = (T *)ar.get("C2T", size_t(m2) * np * sizeof(T));
    d3 = (T *)ar.get("C3", size_t(np) * m3 * sizeof(T));
    if (precision == ZQ_PREC_BF16X3) {
      d1h = (bf16 *)ar.get("C1h", size_t(m1) * ldt * sizeof(bf16));
      d1l = (bf16 *)ar.get("C1l", size_t(m1) * ldt * sizeof(bf16));
    }
    if (!dC || !d1 || !d2t || !d3 || (precision == ZQ_PREC_BF16X3 && (!d1h || !d1l)))
      return zq_fail(ctx, ZQ_ERR_INVALID, "zq_mp2_noalloc: static scratch too small");
    ZQ_TRY(ctx, hipMemcpyAsync(dC, C, size_t(n) * n * 8, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(zq_c_slices<T>, dim3(unsigned(np)), dim3(256), 0, ctx->stream, dC, n, np, m1, m2, m3, d.off3, d1,
                       d2t, d3);
    if (precision == ZQ_PREC_BF16X3)
      hipLaunchKernelGGL(zq_c1_split, dim3(unsigned(m1)), dim3(256), 0, ctx->stream, dC, n, m1, ldt, d1h, d1l);
    ZQ_LAUNCH_CHECK(ctx, "zq_c_slices");
  } else if (cf.dev) {
    // C already on the device: the slices are cut there, as above, into the buffers the host staging
    // below fills
    d1 = (T *)zq_scratch(ctx, "C1", size_t(np) * m1 * sizeof(T), nullptr, false);
    d2t = (T *)zq_scratch(ctx, "C2T", size_t(m2) * np * sizeof(T), nullptr, false);
    d3 = (T *)zq_scratch(ctx, "C3", size_t(np) * m3 * sizeof(T), nullptr, false);
    if (precision == ZQ_PREC_BF16X3) {
      d1h = (bf16 *)zq_scratch(ctx, "C1h", size_t(m1) * ldt * sizeof(bf16), nullptr, false);
      d1l = (bf16 *)zq_scratch(ctx, "C1l", size_t(m1) * ldt * sizeof(bf16), nullptr, false);
    }
    if (!d1 || !d2t || !d3 || (precision == ZQ_PREC_BF16X3 && (!d1h || !d1l)))
      return zq_fail(ctx, ZQ_ERR_OOM, "transform scratch C slices");
    hipLaunchKernelGGL(zq_c_slices<T>, dim3(unsigned(np)), dim3(256), 0, ctx->stream, C, n, np, m1, m2, m3, d.off3, d1,
                       d2t, d3);
    if (precision == ZQ_PREC_BF16X3)
      hipLaunchKernelGGL(zq_c1_split, dim3(unsigned(m1)), dim3(256), 0, ctx->stream, C, n, m1, ldt, d1h, d1l);
    ZQ_LAUNCH_CHECK(ctx, "zq_c_slices");
  } else {
    // coefficient slices, zero rows for padded AOs: C1 (np x m1), C2T (m2 x np), C3 = C4 (np x m3)
    // (each quarter from its own matrix cf.q[0..3]; the same C unless this is a term of the dual transform)
    std::vector<T> h1(size_t(np) * m1, T(0)), h2t(size_t(m2) * np, T(0)), h3(size_t(np) * m3, T(0)),
        h4(own4 ? size_t(np) * m4 : 0, T(0));
    for (int mu = 0; mu < n; ++mu) {
      for (int i = 0; i < m1; ++i) h1[size_t(mu) * m1 + i] = T(cf.q[0][size_t(mu) * n + i]);
      for (int j = 0; j < m2; ++j) h2t[size_t(j) * np + mu] = T(cf.q[1][size_t(mu) * n + j]);
      for (int a = 0; a < m3; ++a) h3[size_t(mu) * m3 + a] = T(cf.q[2][size_t(mu) * n + d.off3 + a]);
      if (own4)
        for (int b = 0; b < m4; ++b) h4[size_t(mu) * m4 + b] = T(cf.q[3][size_t(mu) * n + d.off3 + b]);
    }
    if (int rc = upload(ctx, "C1", h1, &d1)) return rc;
    if (int rc = upload(ctx, "C2T", h2t, &d2t)) return rc;
    if (int rc = upload(ctx, "C3", h3, &d3)) return rc;
    if (own4)
      if (int rc = upload(ctx, "C4", h4, &d4)) return rc;
    if (precision == ZQ_PREC_BF16X3) {
      std::vector<bf16> hh(size_t(m1) * ldt, (bf16)0.f), hl(size_t(m1) * ldt, (bf16)0.f);
      for (int i = 0; i < m1; ++i)
        for (int mu = 0; mu < n; ++mu) {
          const float x = float(C[size_t(mu) * n + i]);
          const bf16 h = (bf16)x;
          hh[size_t(i) * ldt + mu] = h;
          hl[size_t(i) * ldt + mu] = (bf16)(x - (float)h);
        }
      if (int rc = upload(ctx, "C1h", hh, &d1h)) return rc;
      if (int rc = upload(ctx, "C1l", hl, &d1l)) return rc;
    }
    ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));  // the host staging vectors go out of scope below
  }

  if (!d4) d4 = d3;
  const bool sym = cf.q[1] == cf.q[0] && cf.q[3] == cf.q[2] && m1 == m2 && m3 == m4;
  const int nij = sym ? m1 * (m1 + 1) / 2 : m1 * m2;  // batch of Q3 and Q4
  const size_t nX = size_t(Nb) * 64 * m1 * np, nY = size_t(m1) * m2 * np * np, nW = size_t(nij) * np * m3,
               nO = size_t(nij) * m3 * m4;
  T *X = (T *)ar.get("X", nX * sizeof(T));
  T *Y = (T *)ar.get("Y", nY * sizeof(T));
  T *W = (T *)ar.get("W", nW * sizeof(T));
  T *O = (T *)ar.get("OUT", nO * sizeof(T));
  if (!X || !Y || !W || !O) return zq_fail(ctx, fixed ? ZQ_ERR_INVALID : ZQ_ERR_OOM, "transform scratch");

  // Q1: MT (<= 4) tiles of 16 occupied orbitals per workgroup, ceil(m1 / (16 MT)) workgroups along i.
  // Every workgroup along i reads all of its quartets, so one pass (m1 <= 64) beats the higher
  // occupancy of smaller MT (FP32C at MT 4: 1 wave per SIMD, measured 6.1 ms against 8.1 ms in two
  // passes at n = 240, o = 50).
  {
    const int cap = 4;
    const int mt = std::min(cap, (std::min(m1, 16 * cap) + 15) / 16);
    const int groups = precision == ZQ_PREC_BF16X3 ? (m1 + 63) / 64 : (m1 + 16 * mt - 1) / (16 * mt);
    dim3 grid(unsigned(Nb * nb), unsigned(groups));
    const dim3 grid2(unsigned(Nb * nb), unsigned(groups), 2u);  // FP64 / FP32: two halves of the ket slots
    zq_prof_begin(ctx, "ao2mo_q1");
    if (precision == ZQ_PREC_BF16X3) {
      float *Xf = (float *)X;
      switch (mt) {
        case 1: hipLaunchKernelGGL(zq_q1_bf16x3<1>, grid, dim3(512), 0, ctx->stream, ctx->eri, d1h, d1l, ldt, m1, nb, np, Xf); break;
        case 2: hipLaunchKernelGGL(zq_q1_bf16x3<2>, grid, dim3(512), 0, ctx->stream, ctx->eri, d1h, d1l, ldt, m1, nb, np, Xf); break;
        case 3: hipLaunchKernelGGL(zq_q1_bf16x3<3>, grid, dim3(512), 0, ctx->stream, ctx->eri, d1h, d1l, ldt, m1, nb, np, Xf); break;
        default: hipLaunchKernelGGL(zq_q1_bf16x3<4>, grid, dim3(512), 0, ctx->stream, ctx->eri, d1h, d1l, ldt, m1, nb, np, Xf); break;
      }
    } else {
      if constexpr (sizeof(T) == 4) {
        // FP32 forms: read the FP32 copy of the quartets (the same rounding as converting on the fly)
        const float *e32 = ctx->eri32;
        switch (mt) {
          case 1: hipLaunchKernelGGL((zq_q1<MF, 1, float>), grid2, dim3(256), 0, ctx->stream, e32, d1, m1, m1, nb, np, X); break;
          case 2: hipLaunchKernelGGL((zq_q1<MF, 2, float>), grid2, dim3(256), 0, ctx->stream, e32, d1, m1, m1, nb, np, X); break;
          case 3: hipLaunchKernelGGL((zq_q1<MF, 3, float>), grid2, dim3(256), 0, ctx->stream, e32, d1, m1, m1, nb, np, X); break;
          default: hipLaunchKernelGGL((zq_q1<MF, 4, float>), grid2, dim3(256), 0, ctx->stream, e32, d1, m1, m1, nb, np, X); break;
        }
      } else {
        const double *e64 = ctx->eri;
        switch (mt) {
          case 1: hipLaunchKernelGGL((zq_q1<MF, 1, double>), grid2, dim3(256), 0, ctx->stream, e64, d1, m1, m1, nb, np, X); break;
          case 2: hipLaunchKernelGGL((zq_q1<MF, 2, double>), grid2, dim3(256), 0, ctx->stream, e64, d1, m1, m1, nb, np, X); break;
          case 3: hipLaunchKernelGGL((zq_q1<MF, 3, double>), grid2, dim3(256), 0, ctx->stream, e64, d1, m1, m1, nb, np, X); break;
          default: hipLaunchKernelGGL((zq_q1<MF, 4, double>), grid2, dim3(256), 0, ctx->stream, e64, d1, m1, m1, nb, np, X); break;
        }
      }
    }
    zq_prof_end(ctx);
    ZQ_LAUNCH_CHECK(ctx, "ao2mo_q1");
  }
  GemmArgs<T> g2{}, g3{}, g4{};
  // Q2: batch over i; A = C2T (m2 x np), B gathered from X, out Y[i][j][sig][nu] (j <= i if symmetric)
  g2.A = d2t, g2.sAb = 0, g2.sAm = np, g2.sAk = 1;
  g2.B = X, g2.sBb = np, g2.mode = 1, g2.npad = np, g2.sXs = (long long)m1 * np;
  g2.C = Y, g2.sCb = (long long)m2 * np * np, g2.sCm = (long long)np * np, g2.sCn = 1;
  g2.M = m2, g2.N = np * np, g2.K = np;
  g2.tri = sym ? 1 : 0;
  // Q3: batch over (i,j) (the pairs i >= j if symmetric); A = Y_ij [sig][nu], B = C3 [nu][a], out W[ij][sig][a]
  g3.A = Y, g3.sAb = (long long)np * np, g3.sAm = np, g3.sAk = 1;
  g3.B = d3, g3.sBb = 0, g3.sBk = m3, g3.sBn = 1;
  g3.C = W, g3.sCb = (long long)np * m3, g3.sCm = m3, g3.sCn = 1;
  g3.M = np, g3.N = m3, g3.K = np;
  if (sym) g3.tri = 2, g3.sAi = (long long)m2 * np * np, g3.sAj = (long long)np * np;
  // Q4: batch over the same (i,j); A[a][sig] = W[sig][a], B = C4 [sig][b], out[ij][a][b]
  g4.A = W, g4.sAb = (long long)np * m3, g4.sAm = 1, g4.sAk = m3;
  g4.B = d4, g4.sBb = 0, g4.sBk = m4, g4.sBn = 1;
  g4.C = O, g4.sCb = (long long)m3 * m4, g4.sCm = m4, g4.sCn = 1;
  g4.M = m3, g4.N = m4, g4.K = np;
  *packed = sym;
  if constexpr (sizeof(T) == 4) {
    if (precision == ZQ_PREC_BF16X3) {
      if (int rc = launch_gemm_bf16x3(ctx, g2, m1, "ao2mo_q2")) return rc;
      if (int rc = launch_gemm_bf16x3(ctx, g3, nij, "ao2mo_q3")) return rc;
      if (int rc = launch_gemm_bf16x3(ctx, g4, nij, "ao2mo_q4")) return rc;
      *d_out = O;
      return ZQ_OK;
    }
  }
  if (int rc = launch_gemm<MF>(ctx, g2, m1, "ao2mo_q2")) return rc;
  if (int rc = launch_gemm<MF>(ctx, g3, nij, "ao2mo_q3")) return rc;
  // Q4 on the tile kernel in every precision (no library GEMM in the transform)
  if (int rc = launch_gemm<MF>(ctx, g4, nij, "ao2mo_q4")) return rc;
  *d_out = O;
  return ZQ_OK;
}

}  // namespace

// C[m][n] = sum_k A[m][k] B[k][n] (FP64 MFMA, arbitrary element strides), for the SCF driver.
int zq_dgemm(zq_ctx *ctx, const double *A, long long sAm, long long sAk, const double *B, long long sBk, long long sBn,
             double *C, long long sCm, long long sCn, int M, int N, int K, const char *name) {
  GemmArgs<double> g{};
  g.A = A, g.sAm = sAm, g.sAk = sAk;
  g.B = B, g.sBk = sBk, g.sBn = sBn;
  g.C = C, g.sCm = sCm, g.sCn = sCn;
  g.M = M, g.N = N, g.K = K;
  return launch_gemm<MF64>(ctx, g, 1, name);
}

// Runs Q1..Q4 into device memory.  *d_out receives out[i][j][a][b] (dims m1 x m2 x m3 x m4) as double
// (FP64) or float (FP32, BF16X3); *is_f32 tells which.
namespace {
int make_dims(zq_ctx *ctx, int nocc, int block, int precision, Dims &d) {
  if (!ctx->eri) return zq_fail(ctx, ZQ_ERR_STATE, "no ERIs loaded");
  d.n = ctx->n, d.np = ctx->npad, d.nb = ctx->nb, d.Nb = ctx->Nb;
  if (block != ZQ_BLOCK_FULL && block != ZQ_BLOCK_OVOV) return zq_fail(ctx, ZQ_ERR_INVALID, "bad block");
  if (precision != ZQ_PREC_FP64 && precision != ZQ_PREC_FP32 && precision != ZQ_PREC_BF16X3 &&
      precision != ZQ_PREC_FP32C)
    return zq_fail(ctx, ZQ_ERR_INVALID, "bad precision");
  if (block == ZQ_BLOCK_OVOV && (nocc <= 0 || nocc >= d.n)) return zq_fail(ctx, ZQ_ERR_INVALID, "bad nocc");
  d.m1 = block == ZQ_BLOCK_OVOV ? nocc : d.n, d.m2 = d.m1;
  d.off3 = block == ZQ_BLOCK_OVOV ? nocc : 0, d.m3 = d.n - d.off3, d.m4 = d.m3;
  return ZQ_OK;
}

// packed == nullptr: *d_out is the full out[i][j][a][b] (a symmetric chain's pair blocks are expanded);
// else *d_out may hold the pair blocks, as *packed says.
template <class T>
int expand_if(zq_ctx *ctx, const Dims &d, Arena &ar, T *o, bool pk, bool *packed, void **d_out) {
  if (pk && !packed) {
    T *full = (T *)ar.get("OUTF", size_t(d.m1) * d.m2 * d.m3 * d.m4 * sizeof(T));
    if (!full) return zq_fail(ctx, ar.base ? ZQ_ERR_INVALID : ZQ_ERR_OOM, "transform output");
    hipLaunchKernelGGL(zq_expand_pairs<T>, dim3(unsigned(d.m1 * d.m2)), dim3(256), 0, ctx->stream, o, d.m1, d.m3, full);
    ZQ_LAUNCH_CHECK(ctx, "zq_expand_pairs");
    o = full;
  }
  if (packed) *packed = pk;
  *d_out = o;
  return ZQ_OK;
}

int transform_in(zq_ctx *ctx, const Coefs &cf, const Dims &d, int precision, Arena &ar, void **d_out, int dims[4],
                 bool *is_f32, bool *packed = nullptr) {
  dims[0] = d.m1, dims[1] = d.m2, dims[2] = d.m3, dims[3] = d.m4;
  bool pk = false;
  if (precision == ZQ_PREC_FP64) {
    *is_f32 = false;
    double *o = nullptr;
    if (int rc = run_chain<double, MF64>(ctx, cf, d, precision, &o, ar, &pk)) return rc;
    return expand_if(ctx, d, ar, o, pk, packed, d_out);
  }
  *is_f32 = true;
  float *o = nullptr;
  int rc = precision == ZQ_PREC_FP32C ? run_chain<float, MF32C>(ctx, cf, d, precision, &o, ar, &pk)
                                      : run_chain<float, MF32>(ctx, cf, d, precision, &o, ar, &pk);
  if (rc) return rc;
  return expand_if(ctx, d, ar, o, pk, packed, d_out);
}
}  // namespace

int zq_transform_device(zq_ctx *ctx, const double *C, int nocc, int block, int precision, void **d_out, int dims[4],
                        bool *is_f32) {
  Dims d;
  if (int rc = make_dims(ctx, nocc, block, precision, d)) return rc;
  Arena ar{ctx};
  return transform_in(ctx, Coefs{{C, C, C, C}}, d, precision, ar, d_out, dims, is_f32);
}

extern "C" int zq_ao2mo(zq_ctx *ctx, const double *C, size_t nocc, int block, int precision, double *out) {
  ZqGate _gate;
  if (!ctx || !C || !out) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_ao2mo: null argument");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  void *d = nullptr;
  int dims[4];
  bool f32 = false;
  if (int rc = zq_transform_device(ctx, C, int(nocc), block, precision, &d, dims, &f32)) return rc;
  const size_t cnt = size_t(dims[0]) * dims[1] * dims[2] * dims[3];
  const double *src = (const double *)d;
  if (f32) {
    double *wide = (double *)zq_scratch(ctx, "OUT64", cnt * sizeof(double), nullptr, false);
    if (!wide) return zq_fail(ctx, ZQ_ERR_OOM, "zq_ao2mo: widen buffer");
    hipLaunchKernelGGL(zq_widen, dim3(1024), dim3(256), 0, ctx->stream, (const float *)d, cnt, wide);
    ZQ_LAUNCH_CHECK(ctx, "zq_widen");
    src = wide;
  }
  ZQ_TRY(ctx, hipMemcpyAsync(out, src, cnt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ZQ_OK;
}

namespace {
// C and e on the host, or with dev on the device (the context's resident SCF orbitals and energies).
int mp2_in(zq_ctx *ctx, const double *C, const double *e, int nocc, int precision, double *e_corr, Arena &ar,
           bool dev = false) {
  Dims dm;
  if (int rc = make_dims(ctx, nocc, ZQ_BLOCK_OVOV, precision, dm)) return rc;
  if (ar.base && chain_bytes(dm, precision) + 2 * 256 + size_t(ctx->n) * 8 + size_t(nocc) * nocc * 8 + 8 > ar.cap)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_mp2_noalloc: this molecule needs more than the 64 MiB static scratch");
  void *T = nullptr;
  int dims[4];
  bool f32 = false, packed = false;
  if (int rc = transform_in(ctx, Coefs{{C, C, C, C}, dev}, dm, precision, ar, &T, dims, &f32, &packed)) return rc;
  const int o = dims[0], v = dims[2], nij = packed ? o * (o + 1) / 2 : o * o;
  double *de = (double *)ar.get("mp2_e", size_t(ctx->n) * 8);
  double *part = (double *)ar.get("mp2_part", size_t(o) * o * 8 + 8);
  if (!de || !part) return zq_fail(ctx, ar.base ? ZQ_ERR_INVALID : ZQ_ERR_OOM, "mp2 scratch");
  ZQ_TRY(ctx, hipMemcpyAsync(de, e, size_t(ctx->n) * 8, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                             ctx->stream));
  zq_prof_begin(ctx, "mp2_energy");
  if (f32)
    hipLaunchKernelGGL(zq_mp2_pairs<float>, dim3(unsigned(nij)), dim3(256), 0, ctx->stream, (const float *)T, de, o, v,
                       part, int(packed));
  else
    hipLaunchKernelGGL(zq_mp2_pairs<double>, dim3(unsigned(nij)), dim3(256), 0, ctx->stream, (const double *)T, de, o, v,
                       part, int(packed));
  hipLaunchKernelGGL(zq_sum_wave, dim3(1), dim3(64), 0, ctx->stream, part, nij, part + nij);
  zq_prof_end(ctx);
  ZQ_LAUNCH_CHECK(ctx, "zq_mp2_pairs");
  ZQ_TRY(ctx, hipMemcpyAsync(e_corr, part + nij, 8, hipMemcpyDeviceToHost, ctx->stream));
  ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return ZQ_OK;
}

// ---- the static no-allocation scratch (zq_mp2_noalloc)
// kNoAllocBytes hold every intermediate of the MP2 transform for n <= 72 in any precision (the
// molecules of the batched workload: 41.4 MB for (H2O)_3 cc-pVDZ in FP64).  Each device holds its own
// copy for the life of the process.  In host code the array's name is not a device address: its address
// on each device comes from hipGetSymbolAddress, taken once per device with that device current.
constexpr size_t kNoAllocBytes = size_t(64) << 20;

std::mutex g_scratch_mu;
char *g_scratch_ptr[kZqMaxDevices] = {};
int g_scratch_rc[kZqMaxDevices] = {};  // hipError_t of the lookup + 1, 0 before it
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

int zq_mp2_device(zq_ctx *ctx, const double *C, const double *e, int nocc, int precision, double *e_corr) {
  Arena ar{ctx};
  return mp2_in(ctx, C, e, nocc, precision, e_corr, ar);
}

// ---------------------------------------------------------------- grouped MP2 of many molecules
// One launch per stage (C slices, Q1, Q2, Q3, Q4, pair sums, total) over every molecule of a group.
// Workgroup g of a stage belongs to the molecule m with start[m] <= g < start[m + 1] and runs the
// single-molecule body at its local coordinates, so each molecule's arithmetic is exactly zq_mp2's: the
// same kernels on the same shapes (the symmetric chain: Q2 rows j <= i, Q3 and Q4 over the pairs
// i >= j; Q1 with one 16-row tile of occupied orbitals per workgroup, which changes which workgroup
// computes a tile, not how it is summed).
namespace {

template <class T, class E>
struct GJob {
  const E *eri;
  const double *C, *e;
  T *c1, *c2t, *c3, *X, *O;
  double *partial, *out;
  int n, np, nb, Nb, o, v, groups, pairs;
  GemmArgs<T> g[3];  // Q2, Q3, Q4
};

enum { GS_SLICES, GS_Q1, GS_Q2, GS_Q3, GS_Q4, GS_PAIRS, GS_COUNT };

// The molecule whose share of a stage's grid holds workgroup blk: the largest m with start[m] <= blk
// (start[nmol] is the stage's workgroup total, so molecules with no share are passed over).
__device__ inline int group_find(const int *__restrict__ start, int nmol, int blk) {
  int lo = 0, hi = nmol;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= blk)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <class T, class E>
__global__ __launch_bounds__(256) void zq_group_slices(const GJob<T, E> *__restrict__ jobs, const int *__restrict__ start,
                                                       int nmol) {
  const int m = group_find(start, nmol, blockIdx.x);
  const GJob<T, E> J = jobs[m];
  c_slices_body<T>(J.C, J.n, J.np, J.o, J.o, J.v, J.o, J.c1, J.c2t, J.c3, int(blockIdx.x) - start[m]);
}

template <class MF, class E>
__global__ __launch_bounds__(256, (q1_min_waves<MF, 1>())) void zq_group_q1(const GJob<typename MF::T, E> *__restrict__ jobs,
                                                   const int *__restrict__ start, int nmol) {
  const int m = group_find(start, nmol, blockIdx.x);
  const GJob<typename MF::T, E> J = jobs[m];
  const int l = int(blockIdx.x) - start[m], gx = J.Nb * J.nb, t = l / gx;
  q1_body<MF, 1, E>(J.eri, J.c1, J.o, J.o, J.nb, J.np, J.X, l - t * gx, t % J.groups, t / J.groups);
}

template <class MF, class E>
__global__ __launch_bounds__(256, 2) void zq_group_gemm(const GJob<typename MF::T, E> *__restrict__ jobs,
                                                     const int *__restrict__ start, int nmol, int q) {
  const int m = group_find(start, nmol, blockIdx.x);
  const GemmArgs<typename MF::T> g = jobs[m].g[q];
  const int l = int(blockIdx.x) - start[m], gx = (g.N + 63) / 64, gy = (g.M + 63) / 64, t = l / gx;
  gemm_body<MF>(g, l - t * gx, t % gy, t / gy);
}

template <class T, class E>
__global__ __launch_bounds__(256) void zq_group_pairs(const GJob<T, E> *__restrict__ jobs, const int *__restrict__ start,
                                                      int nmol) {
  const int m = group_find(start, nmol, blockIdx.x);
  const GJob<T, E> J = jobs[m];
  mp2_pairs_body<T>(J.O, J.e, J.o, J.v, J.partial, int(blockIdx.x) - start[m], true);
}

template <class T, class E>
__global__ __launch_bounds__(64) void zq_group_sum(const GJob<T, E> *__restrict__ jobs) {
  const GJob<T, E> J = jobs[blockIdx.x];
  sum_wave_body(J.partial, J.pairs, J.out);
}

// Device bytes of one molecule's pieces (each 256-byte aligned) for element type T.
size_t group_bytes(int n, int o, size_t t) {
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t nb = (n + ZQ_BLK - 1) / ZQ_BLK, np = nb * ZQ_BLK, Nb = nb * (nb + 1) / 2, v = n - o, oo = size_t(o) * o,
               pr = size_t(o) * (o + 1) / 2;
  return al(np * o * t) + al(o * np * t) + al(np * v * t) + al(Nb * 64 * o * np * t) + al(oo * np * np * t) +
         al(pr * np * v * t) + al(pr * v * v * t) + al(pr * 8 + 8);
}

constexpr size_t kGroupCap = size_t(8) << 30;  // device bytes of one group's intermediates

template <class T, class MF, class E>
int group_run(hipStream_t stream, const std::vector<ZqGroupMol> &mols, void **ws, size_t *ws_bytes, double *e_corr,
              std::string &err) {
  typedef GJob<T, E> Job;
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const int M = int(mols.size());
  for (int c0 = 0; c0 < M;) {
    // a chunk of consecutive molecules whose pieces fit the cap (at least one molecule)
    int c1 = c0;
    size_t bytes = 0;
    while (c1 < M) {
      const size_t b = group_bytes(mols[c1].n, mols[c1].nocc, sizeof(T));
      if (c1 > c0 && bytes + b > kGroupCap) break;
      bytes += b;
      ++c1;
    }
    const int nm = c1 - c0;
    const size_t head = al(nm * sizeof(Job)) + al(size_t(GS_COUNT) * (nm + 1) * sizeof(int)) + al(nm * 8);
    if (head + bytes > *ws_bytes) {
      if (*ws) (void)hipFree(*ws);
      *ws = nullptr, *ws_bytes = 0;
      if (hipMalloc(ws, head + bytes) != hipSuccess) {
        (void)hipGetLastError();
        *ws = nullptr;
        err = "grouped MP2: out of device memory";
        return ZQ_ERR_OOM;
      }
      *ws_bytes = head + bytes;
    }
    char *base = (char *)*ws;
    Job *d_jobs = (Job *)base;
    int *d_start = (int *)(base + al(nm * sizeof(Job)));
    double *d_out = (double *)(base + al(nm * sizeof(Job)) + al(size_t(GS_COUNT) * (nm + 1) * sizeof(int)));
    size_t off = head;
    auto take = [&](size_t b) {
      char *p = base + off;
      off += al(b);
      return p;
    };
    std::vector<Job> jobs(nm);
    std::vector<int> start(size_t(GS_COUNT) * (nm + 1), 0);
    for (int k = 0; k < nm; ++k) {
      const ZqGroupMol &mo = mols[c0 + k];
      Job &J = jobs[k];
      J.n = mo.n, J.nb = (mo.n + ZQ_BLK - 1) / ZQ_BLK, J.np = J.nb * ZQ_BLK, J.Nb = J.nb * (J.nb + 1) / 2;
      J.o = mo.nocc, J.v = mo.n - mo.nocc, J.groups = (mo.nocc + 15) / 16, J.pairs = mo.nocc * (mo.nocc + 1) / 2;
      const size_t np = J.np, o = J.o, v = J.v, pr = J.pairs;
      if constexpr (std::is_same<E, float>::value)
        J.eri = mo.eri32;
      else
        J.eri = mo.eri;
      J.C = mo.C, J.e = mo.e;
      J.c1 = (T *)take(np * o * sizeof(T));
      J.c2t = (T *)take(o * np * sizeof(T));
      J.c3 = (T *)take(np * v * sizeof(T));
      J.X = (T *)take(size_t(J.Nb) * 64 * o * np * sizeof(T));
      T *Y = (T *)take(o * o * np * np * sizeof(T));
      T *W = (T *)take(pr * np * v * sizeof(T));
      J.O = (T *)take(pr * v * v * sizeof(T));
      J.partial = (double *)take(pr * 8 + 8);
      J.out = d_out + k;
      // Q2: batch over i; A = C2T (o x np), B gathered from X, out Y[i][j][sig][nu], j <= i
      GemmArgs<T> &g2 = J.g[0];
      g2 = GemmArgs<T>{};
      g2.A = J.c2t, g2.sAb = 0, g2.sAm = J.np, g2.sAk = 1;
      g2.B = J.X, g2.sBb = J.np, g2.mode = 1, g2.npad = J.np, g2.sXs = (long long)o * np;
      g2.C = Y, g2.sCb = (long long)o * np * np, g2.sCm = (long long)np * np, g2.sCn = 1;
      g2.M = J.o, g2.N = J.np * J.np, g2.K = J.np, g2.tri = 1;
      // Q3: batch over the pairs i >= j; A = Y_ij [sig][nu], B = C3 [nu][a], out W[p][sig][a]
      GemmArgs<T> &g3 = J.g[1];
      g3 = GemmArgs<T>{};
      g3.A = Y, g3.sAb = (long long)np * np, g3.sAm = J.np, g3.sAk = 1;
      g3.B = J.c3, g3.sBb = 0, g3.sBk = J.v, g3.sBn = 1;
      g3.C = W, g3.sCb = (long long)np * v, g3.sCm = J.v, g3.sCn = 1;
      g3.M = J.np, g3.N = J.v, g3.K = J.np;
      g3.tri = 2, g3.sAi = (long long)o * np * np, g3.sAj = (long long)np * np;
      // Q4: batch over the pairs; A[a][sig] = W[sig][a], B = C3 [sig][b], out[p][a][b]
      GemmArgs<T> &g4 = J.g[2];
      g4 = GemmArgs<T>{};
      g4.A = W, g4.sAb = (long long)np * v, g4.sAm = 1, g4.sAk = J.v;
      g4.B = J.c3, g4.sBb = 0, g4.sBk = J.v, g4.sBn = 1;
      g4.C = J.O, g4.sCb = (long long)v * v, g4.sCm = J.v, g4.sCn = 1;
      g4.M = J.v, g4.N = J.v, g4.K = J.np;
      const long long cnt[GS_COUNT] = {
          J.np,
          (long long)J.Nb * J.nb * J.groups * 2,
          (long long)((g2.N + 63) / 64) * ((g2.M + 63) / 64) * J.o,
          (long long)((g3.N + 63) / 64) * ((g3.M + 63) / 64) * J.pairs,
          (long long)((g4.N + 63) / 64) * ((g4.M + 63) / 64) * J.pairs,
          (long long)J.pairs};
      for (int s = 0; s < GS_COUNT; ++s) {
        const long long next = start[size_t(s) * (nm + 1) + k] + cnt[s];
        if (next > (1LL << 30)) {
          err = "grouped MP2: too many workgroups in one stage";
          return ZQ_ERR_INVALID;
        }
        start[size_t(s) * (nm + 1) + k + 1] = int(next);
      }
    }
    if (hipMemcpyAsync(d_jobs, jobs.data(), nm * sizeof(Job), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipMemcpyAsync(d_start, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, stream) != hipSuccess) {
      err = "grouped MP2: upload of the job table";
      return ZQ_ERR_LAUNCH;
    }
    auto st = [&](int s) { return d_start + size_t(s) * (nm + 1); };
    auto total = [&](int s) { return unsigned(start[size_t(s) * (nm + 1) + nm]); };
    hipLaunchKernelGGL((zq_group_slices<T, E>), dim3(total(GS_SLICES)), dim3(256), 0, stream, d_jobs, st(GS_SLICES), nm);
    hipLaunchKernelGGL((zq_group_q1<MF, E>), dim3(total(GS_Q1)), dim3(256), 0, stream, d_jobs, st(GS_Q1), nm);
    for (int q = 0; q < 3; ++q)
      hipLaunchKernelGGL((zq_group_gemm<MF, E>), dim3(total(GS_Q2 + q)), dim3(256), 0, stream, d_jobs, st(GS_Q2 + q), nm, q);
    hipLaunchKernelGGL((zq_group_pairs<T, E>), dim3(total(GS_PAIRS)), dim3(256), 0, stream, d_jobs, st(GS_PAIRS), nm);
    hipLaunchKernelGGL((zq_group_sum<T, E>), dim3(unsigned(nm)), dim3(64), 0, stream, d_jobs);
    if (hipGetLastError() != hipSuccess) {
      err = "grouped MP2: launch";
      return ZQ_ERR_LAUNCH;
    }
    if (hipMemcpyAsync(e_corr + c0, d_out, nm * 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
      err = "grouped MP2: read-back";
      return ZQ_ERR_LAUNCH;
    }
    c0 = c1;
  }
  return ZQ_OK;
}

}  // namespace

int zq_mp2_grouped(hipStream_t stream, const std::vector<ZqGroupMol> &mols, int precision, void **ws, size_t *ws_bytes,
                   double *e_corr, std::string &err) {
  for (const ZqGroupMol &m : mols)
    if (!m.eri || !m.C || !m.e || m.n <= 0 || m.nocc <= 0 || m.nocc >= m.n ||
        ((precision == ZQ_PREC_FP32 || precision == ZQ_PREC_FP32C) && !m.eri32)) {
      err = "grouped MP2: bad molecule";
      return ZQ_ERR_INVALID;
    }
  if (mols.empty()) return ZQ_OK;
  switch (precision) {
    case ZQ_PREC_FP64: return group_run<double, MF64, double>(stream, mols, ws, ws_bytes, e_corr, err);
    case ZQ_PREC_FP32: return group_run<float, MF32, float>(stream, mols, ws, ws_bytes, e_corr, err);
    case ZQ_PREC_FP32C: return group_run<float, MF32C, float>(stream, mols, ws, ws_bytes, e_corr, err);
    default: err = "grouped MP2: FP64, FP32 or FP32C"; return ZQ_ERR_INVALID;
  }
}

// Called by zq_create with `device` current: the device address of this device's static scratch.
void zq_noalloc_setup(int device) {
  if (device < 0 || device >= kZqMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_rc[device]) return;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_rc[device] = int(e) + 1;
  g_scratch_ptr[device] = e == hipSuccess ? (char *)p : nullptr;
  if (e != hipSuccess) (void)hipGetLastError();
}

extern "C" int zq_mp2(zq_ctx *ctx, const double *C, const double *e, size_t nocc, int precision, double *e_corr) {
  ZqGate _gate;
  if (!ctx || !C || !e || !e_corr) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_mp2: null argument");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  return zq_mp2_device(ctx, C, e, int(nocc), precision, e_corr);
}

extern "C" int zq_mp2_resident(zq_ctx *ctx, size_t nocc, int precision, double *e_corr) {
  ZqGate _gate;
  if (!ctx || !e_corr) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_mp2_resident: null argument");
  if (!ctx->res_C || ctx->res_n != ctx->n || ctx->res_set != ctx->set_id || ctx->res_gen != ctx->eri_gen)
    return zq_fail(ctx, ZQ_ERR_STATE, "zq_mp2_resident: no device SCF result for the ERIs this context holds");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  Arena ar{ctx};
  return mp2_in(ctx, ctx->res_C, ctx->res_e, int(nocc), precision, e_corr, ar, true);
}

extern "C" int zq_mp2_noalloc(zq_ctx *ctx, const double *C, const double *e, size_t nocc, int precision,
                              double *e_corr) {
  ZqGate _gate;
  if (!ctx || !C || !e || !e_corr) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_mp2_noalloc: null argument");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < kZqMaxDevices) base = g_scratch_ptr[ctx->device];
  }
  if (!base) return zq_fail(ctx, ZQ_ERR_LAUNCH, "zq_mp2_noalloc: no static scratch address for this device");
  // profiling would create events: off for this call
  const bool prof = ctx->profile;
  ctx->profile = false;
  Arena ar{ctx, base, kNoAllocBytes, 0};
  const int rc = mp2_in(ctx, C, e, int(nocc), precision, e_corr, ar);
  ctx->profile = prof;
  return rc;
}

// The dual-number transform: value T(g; C,C,C,C) and tangent
// T(dg; C,C,C,C) + T(g; dC,C,C,C) + T(g; C,dC,C,C) + T(g; C,C,dC,C) + T(g; C,C,C,dC), FP64, summed on the
// host in that order.  dg is the quartet set dset holds, read in place of ctx's own for the first term.
extern "C" int zq_ao2mo_dual(zq_ctx *ctx, zq_eri_set *dset, const double *C, const double *dC, size_t nocc, int block,
                             double *out, double *dout) {
  ZqGate _gate;
  if (!ctx || !dset || !C || !dC || !out || !dout) return zq_fail(ctx, ZQ_ERR_INVALID, "zq_ao2mo_dual: null argument");
  if (!ctx->eri) return zq_fail(ctx, ZQ_ERR_STATE, "zq_ao2mo_dual: no ERIs loaded");
  if (dset->n != ctx->n || dset->device != ctx->device)
    return zq_fail(ctx, ZQ_ERR_INVALID, "zq_ao2mo_dual: derivative set of another size or device");
  ZQ_TRY(ctx, hipSetDevice(ctx->device));
  Dims d;
  if (int rc = make_dims(ctx, int(nocc), block, ZQ_PREC_FP64, d)) return rc;
  const size_t cnt = size_t(d.m1) * d.m2 * d.m3 * d.m4;
  std::vector<double> part(cnt);
  const Coefs terms[6] = {{{C, C, C, C}}, {{C, C, C, C}}, {{dC, C, C, C}},
                          {{C, dC, C, C}}, {{C, C, dC, C}}, {{C, C, C, dC}}};
  for (int t = 0; t < 6; ++t) {
    double *dst = t == 0 ? out : (t == 1 ? dout : part.data());
    double *own = ctx->eri;
    if (t == 1) ctx->eri = dset->eri;  // the derivative quartets, same layout
    Arena ar{ctx};
    void *dv = nullptr;
    int dims[4];
    bool f32 = false;
    const int rc = transform_in(ctx, terms[t], d, ZQ_PREC_FP64, ar, &dv, dims, &f32);
    ctx->eri = own;
    if (rc) return rc;
    ZQ_TRY(ctx, hipMemcpyAsync(dst, dv, cnt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ZQ_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (t >= 2)
      for (size_t k = 0; k < cnt; ++k) dout[k] += part[k];
  }
  return ZQ_OK;
}
