// Wave-level FP64 MFMA tile helpers shared by the small dense kernels (zq_dense.hip, zq_group.hip).
//
// tile_nt: one 64-lane wave computes the 16 x 16 tile (i0, j0) of A B^T (A: M x K, B: N x K, both read
// along rows) with v_mfma_f64_16x16x4f64.  Inside each 16-wide K chunk lane l (r = l % 16, q = l / 16)
// holds A[i0+r][k0+4q .. k0+4q+3] and B[j0+r][same k]; the 4 MFMAs of the chunk take element t of every
// lane, so the k order inside a chunk is permuted (the sum is unchanged) and every operand fetch is one
// 32-byte row segment.  The next 64 k are loaded before the current 64 are multiplied.  Accumulator
// element j of lane l is row i0 + 4j + l/16, column j0 + l%16 (the measured gfx950 layout).
#pragma once

#include <hip/hip_runtime.h>

namespace zqt {

typedef double d4 __attribute__((ext_vector_type(4)));

struct Frag {
  double a[4][4], b[4][4];  // [chunk][t]
};

// Loads the 64-k block starting at kb of rows ia (A) and jb (B) for lane group q.
template <bool VEC>
__device__ inline void load_block(Frag &f, const double *__restrict__ A, int lda, const double *__restrict__ B, int ldb,
                                  int ia, int jb, int M, int N, int K, int kb, int q) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int k = kb + 16 * c + 4 * q;
    if (VEC) {
      // K % 16 == 0, even leading dimensions, 16-byte aligned bases: a chunk lies wholly inside K
      const bool kin = kb + 16 * c < K;
      if (kin && ia < M) {
        const double2 *p = reinterpret_cast<const double2 *>(A + size_t(ia) * lda + k);
        const double2 u = p[0], v = p[1];
        f.a[c][0] = u.x, f.a[c][1] = u.y, f.a[c][2] = v.x, f.a[c][3] = v.y;
      } else {
        f.a[c][0] = f.a[c][1] = f.a[c][2] = f.a[c][3] = 0.0;
      }
      if (kin && jb < N) {
        const double2 *p = reinterpret_cast<const double2 *>(B + size_t(jb) * ldb + k);
        const double2 u = p[0], v = p[1];
        f.b[c][0] = u.x, f.b[c][1] = u.y, f.b[c][2] = v.x, f.b[c][3] = v.y;
      } else {
        f.b[c][0] = f.b[c][1] = f.b[c][2] = f.b[c][3] = 0.0;
      }
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        f.a[c][t] = (ia < M && k + t < K) ? A[size_t(ia) * lda + k + t] : 0.0;
        f.b[c][t] = (jb < N && k + t < K) ? B[size_t(jb) * ldb + k + t] : 0.0;
      }
    }
  }
}

// acc = the 16 x 16 tile (i0, j0) of A B^T; element j of lane l is row i0 + 4j + l/16, column j0 + l%16.
template <bool VEC>
__device__ inline d4 tile_nt(const double *__restrict__ A, int lda, const double *__restrict__ B, int ldb, int i0, int j0,
                             int M, int N, int K) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  const int ia = i0 + r, jb = j0 + r;
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  Frag cur, nxt;
  load_block<VEC>(cur, A, lda, B, ldb, ia, jb, M, N, K, 0, q);
  for (int kb = 0; kb < K; kb += 64) {
    if (kb + 64 < K) load_block<VEC>(nxt, A, lda, B, ldb, ia, jb, M, N, K, kb + 64, q);
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a[c][t], cur.b[c][t], acc, 0, 0, 0);
    cur = nxt;
  }
  return acc;
}

// Lane exchanges of a double without LDS: a DPP move inside a 16-lane row, and the gfx950 permlane
// swaps across rows.  A swap of v with itself leaves, in its two results, v and the partner's v (in
// either order, by lane), so their sum is v + v[lane ^ 16] (or ^ 32) on every lane.
template <int CTRL>
__device__ inline double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ inline double add_swap16(double v) {
  const auto lo = __builtin_amdgcn_permlane16_swap(unsigned(__double2loint(v)), unsigned(__double2loint(v)), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(unsigned(__double2hiint(v)), unsigned(__double2hiint(v)), false, false);
  return __hiloint2double(int(hi[0]), int(lo[0])) + __hiloint2double(int(hi[1]), int(lo[1]));
}
__device__ inline double add_swap32(double v) {
  const auto lo = __builtin_amdgcn_permlane32_swap(unsigned(__double2loint(v)), unsigned(__double2loint(v)), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(unsigned(__double2hiint(v)), unsigned(__double2hiint(v)), false, false);
  return __hiloint2double(int(hi[0]), int(lo[0])) + __hiloint2double(int(hi[1]), int(lo[1]));
}

// Sum over the 64 lanes, the same value on every lane: partners at lane distance 1 and 2 (quad
// permutes), then the other quad of the half-row (half-row mirror: the quads are uniform by then), the
// other half of the row (rotate by 8), the other row of the pair (lane ^ 16) and the other half of the
// wave (lane ^ 32).  Every lane adds the same pairs, so every lane holds the same bits.
__device__ inline double wave_sum(double v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x128>(v);
  v = add_swap16(v);
  return add_swap32(v);
}


}  // namespace zqt
