/*
 * zinq_hip.h -- C ABI of the MI355X hot path: J/K Fock build over stored AO ERIs and the AO->MO
 * four-index transform (plus the MP2 energy that consumes it).
 *
 * Plain C99 (size_t / int / double only) so that Zig can pull it in with @cImport next to the
 * reference's other C headers (src/cimport.zig:3-8):
 *     pub const zinqhip = @cImport(@cInclude("zinq_hip.h"));
 *
 * Conventions (those of the reference's libint shim, src/libint.h / src/libint.cpp):
 *   - row-major FP64 arrays; host buffers are owned by the caller, device memory by the context;
 *   - every call is synchronous: results are in the caller's buffer when it returns;
 *   - functions return ZQ_OK (0) or a nonzero status that the Zig wrapper maps to an error
 *     (ZQ_ERR_INVALID -> error.InvalidInput, ZQ_ERR_OOM -> error.GpuOutOfMemory,
 *      ZQ_ERR_LAUNCH -> error.GpuLaunch, ZQ_ERR_NOT_CONVERGED -> error.ScfDidNotConverge,
 *      ZQ_ERR_LAPACK -> error.LapackError); zq_last_error() gives the message;
 *   - one context per host thread; contexts on one GPU run on separate HIP streams.
 * Dense ERIs are in the layout src/libint.cpp:143-150 writes: g[((p*n+q)*n+r)*n+s] = <pq|rs> = (pr|qs).
 */
#ifndef ZINQ_HIP_H
#define ZINQ_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ZQ_OK = 0,
  ZQ_ERR_INVALID = 1,
  ZQ_ERR_OOM = 2,
  ZQ_ERR_LAUNCH = 3,
  ZQ_ERR_NOT_CONVERGED = 4,
  ZQ_ERR_LAPACK = 5,
  ZQ_ERR_STATE = 6
};

/* which integrals zq_ao2mo produces */
enum { ZQ_BLOCK_FULL = 0, ZQ_BLOCK_OVOV = 1 };

/* arithmetic of the quarter transforms; J/K, DIIS, diagonalisation and MP2 sums stay FP64.
 *   ZQ_PREC_FP64    FP64 MFMA, FP64 intermediates
 *   ZQ_PREC_FP32    FP32 MFMA with FP32 accumulation, FP32 intermediates
 *   ZQ_PREC_BF16X3  operands split into two bf16 (hi + lo), hi*hi + hi*lo + lo*hi on bf16 MFMA with FP32
 *                   accumulation (the small products in their own accumulator), FP32 intermediates
 *   ZQ_PREC_FP32C   FP32 MFMA products summed in FP32 over one K block, the blocks summed in FP64
 *                   (compensated sums), FP32 intermediates: the mixed path's transform */
enum { ZQ_PREC_FP64 = 0, ZQ_PREC_FP32 = 1, ZQ_PREC_BF16X3 = 2, ZQ_PREC_FP32C = 3 };

typedef struct zq_ctx zq_ctx;
typedef struct zq_eri_set zq_eri_set; /* resident ERIs shared by contexts, see zq_eri_set_create */

/* Context on HIP device `device` (its own stream).  NULL on failure. */
zq_ctx *zq_create(int device);
void zq_destroy(zq_ctx *ctx);
const char *zq_last_error(const zq_ctx *ctx);

/* Load AO ERIs and keep them resident in HBM, 8-fold packed in 8x8 AO-block quartets.
 * zq_eri_load takes the dense n^4 tensor exactly as libint_coulomb writes it (src/libint.cpp:161-167,
 * consumed by src/molecular_system.zig:115-122); zq_eri_load_packed takes canonical packed chemist ERIs
 * out[I(I+1)/2+J] = (pq|rs), I = p(p+1)/2+q >= J = r(r+1)/2+s (see zinq_ints.h). */
int zq_eri_load(zq_ctx *ctx, const double *g_dense, size_t n);
int zq_eri_load_packed(zq_ctx *ctx, const double *g_packed, size_t n);

/* getFock of src/hartree_fock.zig:686-735 (non-DFT branch), all FP64:
 *   F[k,l] = H[k,l] + sum_ij P[i,j] * (g[i,k,j,l] - x_exch * g[i,j,k,l]),  then F <- (F + F^T)/2,
 * with x_exch = 0.5 for RHF and 1.0 for GHF (spin-blocked ERIs loaded).  H, P, F_out are n x n. */
int zq_fock_jk(zq_ctx *ctx, const double *H, const double *P, double x_exch, double *F_out);

/* getFock for ndens densities in one pass over the resident ERIs (the densities of a work item run back to
 * back and share its quartet reads): F_out[d] = H + G(P[d]), each symmetrised, P and F_out ndens x n x n.
 * H may be NULL, giving the bare two-electron part G(P[d]): the response contraction
 * V_x = sum P_x (g - x g) of CPHF (src/cphf.zig:167-172).  The DFT branch's J + exx K
 * (src/hartree_fock.zig:701-719) is zq_fock_jk with x_exch = exx / 2. */
int zq_fock_jk_multi(zq_ctx *ctx, const double *H, const double *P, size_t ndens, double x_exch, double *F_out);

/* The two-electron part of the HF nuclear gradient, g_G of src/hartree_fock.zig:263-273: for each of
 * ncoord derivative ERI sets (dg_packed: ncoord consecutive canonical packed arrays of the
 * zq_eri_load_packed layout, each the derivative of (pq|rs) with respect to one nuclear coordinate),
 *   out[x] = sum_pqrs 1/2 P[p,q] P[r,s] ((pq|rs)_x - x_exch (pr|qs)_x),  x_exch = exx / 2 (RHF; 1/2 for HF), exx (GHF),
 * one J/K pass over each set on the GPU (a private context holds the set; ctx's own ERIs stay loaded). */
int zq_grad_eri(zq_ctx *ctx, const double *dg_packed, size_t n, size_t ncoord, const double *P, double x_exch,
                double *out);

/* AO->MO transform of the loaded ERIs with MO coefficients C (n x n, columns = MOs).
 *   ZQ_BLOCK_FULL: out[p,q,r,s] = <pq|rs> for all MOs, the result of ao2mo_pppp
 *                  (src/integral_transform.zig:37-79); out has n^4 entries.
 *   ZQ_BLOCK_OVOV: out[i,j,a,b] = <ij|ab> = (ia|jb), i,j < nocc, a,b >= nocc, the result of ao2mo_oovv
 *                  (src/integral_transform.zig:165-219); out has nocc^2 (n-nocc)^2 entries.
 * precision: ZQ_PREC_FP64 | ZQ_PREC_FP32 | ZQ_PREC_BF16X3 | ZQ_PREC_FP32C (quarter transforms only). */
int zq_ao2mo(zq_ctx *ctx, const double *C, size_t nocc, int block, int precision, double *out);

/* The dual-number transform of mp's gradient path (src/moller_plesset.zig:304-333, ao2mo_pppp over
 * ScalarDual values): the transform of the ERI dual g + eps dg with the coefficient dual C + eps dC,
 *   out  = T(g; C, C, C, C),
 *   dout = T(dg; C, C, C, C) + T(g; dC, C, C, C) + T(g; C, dC, C, C) + T(g; C, C, dC, C) + T(g; C, C, C, dC),
 * T(x; A, B, A', B')[i,j,a,b] = sum A[mu,i] B[lam,j] A'[nu,a] B'[sig,b] <mu lam|nu sig>_x, in FP64 on the GPU.
 * g is ctx's resident ERIs; dg the quartets of dset (an ERI set of the same size, e.g. the derivative with
 * respect to one nuclear coordinate).  block, nocc and the output layout are those of zq_ao2mo. */
int zq_ao2mo_dual(zq_ctx *ctx, zq_eri_set *dset, const double *C, const double *dC, size_t nocc, int block,
                  double *out, double *dout);

/* Closed-shell MP2 correlation energy from the OVOV transform, FP64 denominators and summation:
 *   E2 = sum_ijab (ia|jb) [2 (ia|jb) - (ib|ja)] / (e_i + e_j - e_a - e_b),
 * equal to the order-2 term of mp() (src/moller_plesset.zig:340-481) for an RHF reference. */
int zq_mp2(zq_ctx *ctx, const double *C, const double *e, size_t nocc, int precision, double *e_corr);

/* zq_mp2 on the orbitals and orbital energies of the last successful zq_scf_run_device on ctx
 * (include/zinq_scf.h), which stay on the device: the same kernels and result as zq_mp2 with the C and
 * e that call returned, without sending them back to the device.  ZQ_ERR_STATE if ctx holds no such
 * result for its current ERIs (none yet, a failed run, or zq_eri_load* / zq_attach to another set
 * since).  No counterpart in the reference, which hands the orbitals to mp() in host memory: it is for
 * request loops that keep a molecule's ERIs and MO coefficients resident. */
int zq_mp2_resident(zq_ctx *ctx, size_t nocc, int precision, double *e_corr);

/* ---- determinant consumers of the FULL transform ----
 * Both run the FULL transform of the resident ERIs (FP64) on the GPU and read the spin-orbital
 * integrals from it on the fly: spin orbital p = 2P + s for a restricted reference (the column order of
 * ao2so_coef, src/integral_transform.zig:82-101; spatial ERIs loaded, C n x n), or the loaded spin-blocked
 * ERIs and a 2 nbf x 2 nbf C for generalized = 1.  Determinants (at most 128 spin orbitals) are those of
 * generateDets (src/configuration_interaction.zig:126-183); H_CI is built on the GPU with the
 * Slater-Condon rules of slater (:323-432), dense, at most 16384 determinants.  *ndet_out (may be NULL)
 * receives the determinant count.
 *
 * zq_ao2mo_pp: out = C^T A C for n x n A and C, ao2mo_pp (src/integral_transform.zig:12-34).
 * zq_ci: computeCiStates (src/configuration_interaction.zig:597-683): the lowest min(nstates, ndet)
 *   eigenvalues of H_CI over the excitation levels excitations[0..nexc) (1 = CIS, {1, 2} = CISD), + VN;
 *   H is the core Hamiltonian in the basis of C (spin-blocked for generalized).
 * zq_mp: mp (src/moller_plesset.zig:340-481): energies[0..order] of Rayleigh-Schroedinger perturbation
 *   theory over singles .. order-fold excitations with H_MS = diag(e) - sum_i (<pi|qi> - <pi|iq>);
 *   energies[0] is the sum of the occupied orbital energies, the total energy is E_HF + sum_{k>=2}.
 *   Order 2 forms only row 0 of H_CI (up to 2^24 determinants); higher orders the dense H_CI. */
int zq_ao2mo_pp(zq_ctx *ctx, size_t n, const double *A, const double *C, double *out);
int zq_ci(zq_ctx *ctx, const double *H, const double *C, size_t nocc, int generalized, const int *excitations,
          size_t nexc, double VN, size_t nstates, double *energies, size_t *ndet_out);
int zq_mp(zq_ctx *ctx, const double *C, const double *e, size_t nocc, int generalized, int order, double *energies,
          size_t *ndet_out);

/* zq_mp2 that allocates nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc).  Same contract,
 * kernels and result as zq_mp2, but every intermediate lives in one static device array of 64 MiB
 * per GPU, held for the life of the process: molecules up to n = 72 fit (n x nocc dependent; a larger
 * one returns ZQ_ERR_INVALID with the reason).  The FP32 forms need the FP32 ERI copy already resident
 * (an ERI set created with_fp32, or an earlier FP32 call on ctx), else ZQ_ERR_STATE.  All calls on one
 * GPU share that array: calls that overlap in time on one GPU (several threads or contexts) race and
 * must be serialised by the caller. */
int zq_mp2_noalloc(zq_ctx *ctx, const double *C, const double *e, size_t nocc, int precision, double *e_corr);

/* ---- measurement hooks (no counterpart in the reference) ----
 * With profiling on, every kernel launch is bracketed by HIP events on the context's stream;
 * zq_profile_get returns the summed milliseconds and launch count of kernels whose name starts
 * with `prefix` since the last zq_profile_reset. */
void zq_profile_enable(zq_ctx *ctx, int on);
void zq_profile_reset(zq_ctx *ctx);
int zq_profile_get(zq_ctx *ctx, const char *prefix, double *ms, int *launches);

/* Device memory held by the context, in bytes (ERIs + scratch). */
size_t zq_device_bytes(const zq_ctx *ctx);

/* Resident ERI sets for serving: the quartets of one molecule (and, with with_fp32, their FP32 copy for
 * the mixed Fock path) stay in HBM independently of any context, together with the molecule's J/K
 * work plan, and several contexts -- one per concurrent request, each with its own stream, J/K
 * records and scratch -- compute on them at once.  A context's buffers only grow, so re-attaching
 * allocates nothing once they fit the largest molecule.
 * zq_eri_set_create takes the zq_eri_load_packed layout; NULL on failure.  zq_eri_set_bytes is the
 * set's HBM footprint (the quartets and their FP32 copy; the plan adds under 1%).  zq_attach points
 * ctx at the set (no copy; the set must outlive the attachment,
 * which ends with the next zq_attach, zq_eri_load* or zq_destroy of ctx). */
zq_eri_set *zq_eri_set_create(int device, const double *g_packed, size_t n, int with_fp32);
void zq_eri_set_destroy(zq_eri_set *s);
size_t zq_eri_set_bytes(const zq_eri_set *s);
int zq_attach(zq_ctx *ctx, zq_eri_set *s);

/* J/K records (64 doubles each) a context keeps for a molecule of n basis functions: the size of its
 * grow-only record buffer, for admission control (host arithmetic, no device needed). */
size_t zq_jk_record_count(size_t n);

/* Free and total HBM of `device` in bytes (for admission control). */
int zq_mem_info(int device, size_t *free_bytes, size_t *total_bytes);

#ifdef __cplusplus
}
#endif

#endif /* ZINQ_HIP_H */
