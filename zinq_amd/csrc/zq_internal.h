// Internal declarations shared by the HIP translation units of libzinq_hip.so.
//
// Device ERI layout ("block quartets").  AO indices are grouped in blocks of 8 (n padded to npad =
// 8*nb with zero functions).  A block pair (bp >= bq) has index bpair(bp,bq) = bp(bp+1)/2 + bq; there
// are Nb = nb(nb+1)/2 of them.  Every unique integral (pq|rs) lives in exactly one block quartet
// (Ib >= Jb) of 64 x 64 doubles, element [ket local r*8+s][bra local p*8+q] (bra fastest).  Positions
// that duplicate another element of the same quartet (p < q when bp == bq, r < s when br == bs, and bra
// local < ket local when Ib == Jb) hold 0.  Quartets are stored by Ib, and for each Ib its Ib+1 ket
// blocks Jb <= Ib in the order bs ascending, then br ascending, which is the order the J/K kernel walks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/zinq_hip.h"

#define ZQ_BLK 8
#define ZQ_QUARTET 4096  // 64 x 64 doubles

__host__ __device__ inline int zq_bpair(int a, int b) { return a * (a + 1) / 2 + b; }

// Number of ket blocks in the row of bra block pair (bp,bq) with ket column bs.
__host__ __device__ inline int zq_run_len(int bp, int bq, int bs) { return (bs <= bq ? bp : bp - 1) - bs + 1; }

// Position of ket block pair (br,bs) in the ket list of bra block pair (bp,bq); requires
// bpair(br,bs) <= bpair(bp,bq).
__host__ __device__ inline int zq_ket_pos(int bp, int bq, int br, int bs) {
  const int extra = bs - bq - 1;
  return bs * (bp + 1) - bs * (bs - 1) / 2 - (extra > 0 ? extra : 0) + (br - bs);
}

__host__ __device__ inline size_t zq_blk_index(int bp, int bq, int br, int bs) {
  const size_t Ib = size_t(zq_bpair(bp, bq));
  return Ib * (Ib + 1) / 2 + size_t(zq_ket_pos(bp, bq, br, bs));
}

// Inverse of zq_bpair for small indices.
__host__ __device__ inline void zq_unpair(int I, int &a, int &b) {
  a = 0;
  while ((a + 1) * (a + 2) / 2 <= I) ++a;
  b = I - a * (a + 1) / 2;
}

struct JKItem {
  int Ib, bp, bq, k0, k1, rec0;
  int mol;  // molecule of a grouped plan (0 for a single molecule)
};

// Per-molecule table of a grouped J/K build: quartets at eri + eri_off, P at P + p_off, H at
// H + h_off, F at F + f_off, the three matrices with leading dimension ld (zero padded beyond n).
struct JKMol {
  long long eri_off, p_off, h_off, f_off;
  int ld, n;
};

// Work plan of one molecule's J/K build: items, the number of 64-double records they emit, and the CSR
// list (per output block x >= y) of record codes r << 1 | transposed.
struct JKPlan {
  std::vector<JKItem> items;
  std::vector<int> ptr, idx;
  size_t nrecords = 0;
};
// chunk = quartets per item (<= 0: the default for this size)
void zq_jk_plan(int Nb, size_t nquartets, int chunk, JKPlan &plan);  // zq_jk.hip
// The same plan on the device: it depends on the molecule's size only, so a resident ERI set keeps one
// that every context attached to it reads.
struct JKPlanDev {
  JKItem *items = nullptr;
  int *ptr = nullptr, *idx = nullptr;
  int nitems = 0;
  size_t nrecords = 0;
};
// Builds the default plan for Nb block pairs and uploads it on st (synchronised before returning);
// ZQ_OK or an error code with its message in err.  zq_jk_plan_free releases the three arrays.
int zq_jk_plan_upload(int Nb, size_t nquartets, hipStream_t st, JKPlanDev &d, std::string &err);  // zq_jk.hip
void zq_jk_plan_free(JKPlanDev &d);                                                             // zq_jk.hip

struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
};

struct zq_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // ERIs
  int n = 0, npad = 0, nb = 0, Nb = 0;
  size_t nquartets = 0;
  double *eri = nullptr;
  float *eri32 = nullptr;  // FP32 copy for the mixed-precision Fock builds (zq_eri32), same layout
  bool eri_owned = true, eri32_owned = true;  // false while attached to a shared zq_eri_set

  // J/K work decomposition: the plan (items, CSR) is the context's own after zq_eri_load and the set's
  // while attached; records, P, H and F are the context's, grow-only (capacities in doubles), so that
  // re-attaching frees nothing once they fit the largest molecule
  JKItem *d_items = nullptr;
  int nitems = 0;
  size_t nrecords = 0;
  double *d_records = nullptr;
  int *d_csr_ptr = nullptr, *d_csr_idx = nullptr;
  bool plan_owned = true;
  unsigned long long set_id = 0;  // id of the attached zq_eri_set, 0 when none
  double *d_P = nullptr, *d_H = nullptr, *d_F = nullptr;
  size_t cap_records = 0, cap_P = 0, cap_H = 0, cap_F = 0;

  // rocBLAS handle used by the device SCF's eigensolver (created on first use)
  void *solver = nullptr;
  // fork/join events of the eigensolver's Q accumulation, which runs on a stream shared by the device's
  // contexts beside the tridiagonal solve
  hipEvent_t eig_events[2] = {nullptr, nullptr};

  // captured SCF iterations (hipGraphExec_t) by the key they were captured for: sizes, options, and
  // the residency (the attached set's id, else the ERI load's generation), so that a context serving
  // several molecules keeps one per molecule.  scratch_gen changes whenever a buffer a capture holds
  // is re-allocated; the cache is emptied then (scf_graphs_gen records the generation it holds).
  std::map<std::string, void *> scf_graphs;
  unsigned scf_graphs_gen = 0;
  unsigned scf_captures = 0;  // SCF iterations captured over the context's lifetime (diagnostics)
  unsigned scratch_gen = 0, eri_gen = 0;
  int sp2_budget = 100;  // SP2 steps per density in the asynchronous SCF; grows after a fallback
  int sp2_guess_budget = 128;  // SP2 steps for the core-Hamiltonian guess density of that SCF
  // the orbitals (n x n, C[mu][k]) and orbital energies of the last successful zq_scf_run_device, left on
  // the device for zq_mp2_resident, with the residency they belong to (null: none)
  const double *res_C = nullptr, *res_e = nullptr;
  int res_n = 0;
  unsigned long long res_set = 0;
  unsigned res_gen = 0;
  double diis_err_scale = 1.0;  // factor on the device SCF's stored DIIS error vectors (a test hook)
  int scf_diis_resets = 0;  // singular DIIS systems the last device SCF reset its history on
  int scf_replays = 0;   // iterations the asynchronous driver enqueued in the last device SCF
  int scf_driver = 0;    // driver that finished the last device SCF: 1 device-decided, 2 host-synchronised
  int scf_sp2[3] = {0, 0, 0};  // its SP2 densities, folded ones, steps (device-decided driver only)
  hipEvent_t scf_events[2] = {nullptr, nullptr};  // the asynchronous SCF's lagged state reads

  // named scratch buffers (grow-only)
  std::map<std::string, DevBuf> scratch;
  // named pinned host buffers (hipHostMalloc, grow-only) for asynchronous device-to-host reads
  std::map<std::string, DevBuf> pinned;

  // profiling
  bool profile = false;
  struct Ev {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Ev> events;
  std::map<std::string, std::pair<double, int>> prof;
};

// Resident ERIs of one molecule, shared read-only by the contexts attached to it (zq_attach).
struct zq_eri_set {
  int device = 0;
  int n = 0, nb = 0, npad = 0, Nb = 0;
  size_t nquartets = 0;
  double *eri = nullptr;
  float *eri32 = nullptr;
  JKPlanDev plan;  // the J/K plan of this size, shared by the attached contexts
  unsigned long long id = 0;  // unique per process: a new set may reuse a freed set's addresses
  std::string err;
};

// ---- process-wide gate between hipGraph stream captures and the runtime calls that break them
// (allocation, free, synchronous copies, rocSOLVER/rocBLAS) in other threads.  Every public entry point
// holds a ZqGate (shared, re-entrant per thread); a capture runs inside a ZqCaptureWindow (exclusive;
// the thread's own shared hold is released for the window and taken back after it).
struct ZqGate {
  ZqGate();
  ~ZqGate();
  ZqGate(const ZqGate &) = delete;
  ZqGate &operator=(const ZqGate &) = delete;
};
struct ZqCaptureWindow {
  int saved;
  ZqCaptureWindow();
  ~ZqCaptureWindow();
  ZqCaptureWindow(const ZqCaptureWindow &) = delete;
  ZqCaptureWindow &operator=(const ZqCaptureWindow &) = delete;
};

// ---- helpers implemented in zq_context.hip
int zq_fail(zq_ctx *ctx, int code, const std::string &msg);
int zq_check(zq_ctx *ctx, hipError_t e, const char *what);
// Named grow-only device buffer of the context.  `captured`: the buffer may be held by a captured SCF
// iteration, so re-allocating it empties the capture cache (transform and multi-density buffers, never
// captured, pass false).
void *zq_scratch(zq_ctx *ctx, const std::string &name, size_t bytes, bool *fresh = nullptr, bool captured = true);
void *zq_pinned(zq_ctx *ctx, const std::string &name, size_t bytes);
void zq_free_eri(zq_ctx *ctx);
void zq_prof_begin(zq_ctx *ctx, const char *name);
void zq_prof_end(zq_ctx *ctx);
void zq_release_solver(zq_ctx *ctx);  // zq_scf_dev.hip, also drops the captured SCF graph
// The context's rocBLAS handle (rocblas_handle), bound to its stream, atomics off; null on failure.
void *zq_blas(zq_ctx *ctx);  // zq_eigen.hip
int zq_fock_jk_device(zq_ctx *ctx, const double *dP, const double *dH, double x_exch, double *dF,
                      const int *skip = nullptr, bool f32 = false, double *epart = nullptr);  // zq_jk.hip
int zq_eri32(zq_ctx *ctx);  // zq_jk.hip: FP32 copy of the resident quartets, made once per load
void zq_narrow_eri(hipStream_t stream, const double *eri, size_t cnt, float *out);  // zq_jk.hip
// MP2 correlation energy of the resident ERIs (zq_transform.hip); C and e are host arrays.
int zq_mp2_device(zq_ctx *ctx, const double *C, const double *e, int nocc, int precision, double *e_corr);
// One molecule of a grouped MP2: its resident quartets (and their FP32 copy, for the FP32 forms), size,
// doubly occupied count, and device C (n x n, row-major, columns = MOs) and orbital energies e (n).
struct ZqGroupMol {
  const double *eri;
  const float *eri32;
  int n, nocc;
  const double *C, *e;
};
// Closed-shell MP2 correlation energies of many molecules, one launch per transform stage for all of
// them, each molecule's result bitwise that of zq_mp2 (zq_transform.hip).  precision: FP64, FP32 or
// FP32C.  ws / ws_bytes: the caller's grow-only device workspace (allocated here when too small, under
// the caller's gate).  e_corr: host array of mols.size().  On error the reason is in err.
int zq_mp2_grouped(hipStream_t stream, const std::vector<ZqGroupMol> &mols, int precision, void **ws, size_t *ws_bytes,
                   double *e_corr, std::string &err);
// Q1..Q4 of the resident ERIs into context scratch (zq_transform.hip): *d_out = out[i][j][a][b] with dims
// m1 x m2 x m3 x m4 (FULL: n^4, <pq|rs>; OVOV: o o v v), double or (*is_f32) float.
int zq_transform_device(zq_ctx *ctx, const double *C, int nocc, int block, int precision, void **d_out, int dims[4],
                        bool *is_f32);
// Devices whose static no-allocation scratch address is kept (zq_transform.hip).
constexpr int kZqMaxDevices = 64;
// Looks up, once per device, the device address of zq_mp2_noalloc's static scratch; called by zq_create
// with `device` current (zq_transform.hip).
void zq_noalloc_setup(int device);
// Grouped J/K on `stream`: F = H + G(P) (x_exch) for every molecule of a grouped plan whose done flag
// is 0.  blk[g] = (molecule, output block) for the nblk CSR rows.  Returns the launch status.
hipError_t zq_jk_group(hipStream_t stream, const double *eri, const JKItem *items, int nitems, const double *P,
                       double x_exch, double *rec, const JKMol *mols, const int *done, const int *csr_ptr,
                       const int *csr_idx, const int2 *blk, int nblk, const double *H, double *F);
// Packs canonical chemist ERIs (zq_eri_load_packed layout) into 8x8 AO-block quartets (zq_context.hip).
__global__ __launch_bounds__(256) void zq_pack_quartets(const double *__restrict__ canon, double *__restrict__ out, int n,
                                                        size_t nquartets);
int zq_dgemm(zq_ctx *ctx, const double *A, long long sAm, long long sAk, const double *B, long long sBk, long long sBn,
             double *C, long long sCm, long long sCn, int M, int N, int K, const char *name);  // zq_transform.hip
// C = alpha A B^T (A: M x K, B: N x K, row-major with leading dimensions), latency-sized FP64 MFMA GEMM.
// With bsel != nullptr, B is offset by *bsel * bstride elements on the device (choice made by a kernel).
int zq_gemm_nt(zq_ctx *ctx, const double *A, int lda, const double *B, int ldb, double *C, int ldc, int M, int N, int K,
               double alpha, const char *name, const int *skip = nullptr, const int *bsel = nullptr,
               long long bstride = 0);  // zq_dense.hip
// SP2 purification of the symmetric n x n dF (ld n): *dD = the projector onto its nocc lowest
// eigenvectors (m x m, m = n rounded up to 16, zero padded, *ld = m, in context scratch).  ZQ_OK, or
// ZQ_ERR_NOT_CONVERGED when max_steps did not reach idempotency (no gap); *steps_hint is the number of
// steps before the first convergence check and is updated on success.
int zq_sp2_density(zq_ctx *ctx, const double *dF, int n, int nocc, double tol, int max_steps, int *steps_hint,
                   const double **dD, int *ld, int *steps);  // zq_dense.hip
// Enqueues SP2 with a budget of nsteps <= kSp2MaxSteps steps (one launch that stops once settled) and no
// host synchronisation.  state6 = two {count, steps} pairs, the parity of the buffer that holds the
// settled projector, and an abort flag: the projector is *dX + state6[4] * *stride (m x m, ld m), which
// a kernel selects at run time.  The final pair is state6 + 2 * (nsteps & 1); the result is settled when
// count >= 2 there and state6[5] == 0.  All kernels return at once while *skip != 0.  track = 1 (the
// first density of a device SCF) or 2 (a later one of the same SCF) lets the run fold with HOMO/LUMO
// bounds certified from an earlier plain run (zq_dense.hip); 0 runs plain SP2 only.
constexpr int kSp2MaxSteps = 128;
// Eigen-decomposition of the symmetric n x n A (ld n, overwritten): w = eigenvalues ascending, row k of
// A = eigenvector k.  offd: n doubles of workspace; info: device int (0 on success).  n <= 192 by the
// one-workgroup tridiagonal reduction of zq_eigen.hip and rocSOLVER dstedc, larger n by dsyevd.
int zq_syev_rows(zq_ctx *ctx, double *A, int n, double *w, double *offd, int *info);  // zq_eigen.hip
int zq_sp2_enqueue(zq_ctx *ctx, const double *dF, int n, int nocc, double tol, int nsteps, const int *skip, int *state6,
                   const double **dX, long long *stride, int *ld, int track);  // zq_dense.hip
// {densities, folded densities, SP2 steps} of the tracked runs since the last track = 1 run (synchronous).
int zq_sp2_tracked_stats(zq_ctx *ctx, int out[3]);  // zq_dense.hip

#define ZQ_TRY(ctx, call)                                \
  do {                                                   \
    hipError_t _e = (call);                              \
    if (_e != hipSuccess) return zq_check(ctx, _e, #call); \
  } while (0)

#define ZQ_LAUNCH_CHECK(ctx, name)                                                   \
  do {                                                                               \
    hipError_t _e = hipGetLastError();                                               \
    if (_e != hipSuccess) return zq_fail(ctx, ZQ_ERR_LAUNCH, std::string(name) + ": " + hipGetErrorString(_e)); \
  } while (0)
