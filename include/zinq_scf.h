/*
 * zinq_scf.h -- host SCF driver over the zinq_hip.h hot path.
 *
 * A C++ restatement of the reference's Zig host loop, src/hartree_fock.zig:738-875 (scf) preceded by
 * the core-Hamiltonian guess of :425-431, in which getFock (:686-735) is the GPU call zq_fock_jk and
 * the rest (getEnergy :643-655, getError :658-683, diis :176-227, geigh -> dsygvd, getDensity :621-640)
 * runs in FP64 on the host with in-house linear algebra.  The ERIs must already be resident in `ctx`
 * (zq_eri_load / zq_eri_load_packed): spatial for RHF, spin-blocked (coulombSpin) for GHF.
 */
#ifndef ZINQ_SCF_H
#define ZINQ_SCF_H

#include <stddef.h>

#include "zinq_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int diis;          /* DIIS history length; 0 disables DIIS (Options.diis = null), default 8 */
  int iterations;    /* default 100 */
  double threshold;  /* on |dE| and rms(dP), default 1e-8 */
  int generalized;   /* 0: RHF (x = 1/2, P = 2 C C^T); 1: GHF on spin-blocked integrals (x = 1) */
  int fock_precision; /* zq_scf_run_device only.  ZQ_PREC_FP64: every Fock build in FP64 (the reference's
                         arithmetic).  ZQ_PREC_FP32: the mixed path -- incremental builds
                         F_k = F_(k-1) + G(P_k - P_(k-1)) from an FP32 copy of the ERIs (FP64 products
                         and sums) until convergence, then one full FP64 build from the last density for
                         the energy and the orbitals.  zq_scf_run always uses FP64. */
} zq_scf_options;

typedef struct {
  double energy;   /* total energy of the last iteration, E_elec + V_NN */
  int iterations;  /* iterations run */
} zq_scf_result;

/* H, S: n x n (n = 2 nbf for GHF).  nocc: doubly occupied orbitals (RHF) or electrons (GHF).
 * Outputs (caller-allocated, may not be NULL): C (n x n, MOs in columns), e (n), P (n x n),
 * F (n x n, the last diagonalised, DIIS-extrapolated Fock matrix).
 * Returns ZQ_OK, ZQ_ERR_NOT_CONVERGED (error.ScfDidNotConverge), ZQ_ERR_LAPACK, or a GPU error. */
int zq_scf_run(zq_ctx *ctx, size_t n, size_t nocc, const double *H, const double *S, double VN,
               const zq_scf_options *opt, double *C, double *e, double *P, double *F, zq_scf_result *res);

/* Same contract as zq_scf_run with every n x n step on the GPU.  Each iteration is decided on the
 * device: J/K, energy, DIIS error, the DIIS history and its bordered LU solve, the extrapolation, and
 * the density P = f X D X^T (X = L^-T from S = L L^T), where D, the projector onto the nocc lowest
 * eigenvectors of X^T F X, comes from SP2 purification (GEMMs on MFMA).  The iteration is captured as a
 * hipGraph (one per molecule and options, kept by ctx) and replayed; the host reads the convergence state one iteration late, so the GPU is
 * never idle between iterations.  After convergence one zq_eigh decomposition of the last
 * extrapolated F (in the orthonormal basis) gives C and e.  A density whose SP2 does not settle within the step budget (no HOMO-LUMO gap) makes
 * the run restart on a host-synchronised driver that falls back to dsyevd for that density. */
int zq_scf_run_device(zq_ctx *ctx, size_t n, size_t nocc, const double *H, const double *S, double VN,
                      const zq_scf_options *opt, double *C, double *e, double *P, double *F, zq_scf_result *res);

/* SP2 steps per density in zq_scf_run_device's device-decided iteration (1..128, default 100).  SP2 is
 * one launch that stops as soon as the density has settled, so the budget only bounds it.  A density
 * that needs more restarts the run on the host-synchronised driver and raises the budget by 16. */
int zq_scf_set_sp2_budget(zq_ctx *ctx, int steps);

/* The same for the core-Hamiltonian guess density of that iteration (1..128, default 128).  A guess
 * that does not settle ends the device-decided run after its first iteration and restarts on the
 * host-synchronised driver; unlike the per-iteration budget it is not raised.  A tuning and test hook
 * with no counterpart in the reference. */
int zq_scf_set_sp2_guess_budget(zq_ctx *ctx, int steps);

/* A test hook with no counterpart in the reference: zq_scf_run_device stores its DIIS error vectors
 * FPS - SPF (src/hartree_fock.zig:658-683) multiplied by scale, in [0, 1] (default 1, which changes no
 * bit).  At 0 every DIIS system from the second iteration on is singular, so each iteration takes the
 * reset of src/hartree_fock.zig:821-845 (history cut to the current entry, F left unextrapolated) and
 * the SCF is the one without DIIS.  The convergence test (|dE|, rms(dP)) does not use the errors. */
int zq_scf_set_diis_error_scale(zq_ctx *ctx, double scale);

/* Diagnostics, with no counterpart in the reference: how many singular DIIS systems the last
 * zq_scf_run_device on ctx met (each one resets the history to the current entry), from the driver
 * that finished the run. */
int zq_scf_last_diis_resets(const zq_ctx *ctx);

/* Diagnostics, with no counterpart in the reference: the iterations the device-decided driver of the
 * last zq_scf_run_device on ctx enqueued (graph replays; the host reads the state one iteration late,
 * so a run that ends after iteration k enqueued k + 1 or k + 2 of them), 0 if it did not run. */
int zq_scf_last_replays(const zq_ctx *ctx);

/* Diagnostics, with no counterpart in the reference: which driver finished the last zq_scf_run_device
 * on ctx -- 1 the device-decided iteration, 2 the host-synchronised restart -- or 0 before any run. */
int zq_scf_last_driver(const zq_ctx *ctx);

/* Diagnostics, with no counterpart in the reference: how many SCF iterations zq_scf_run_device has
 * captured as hipGraphs on ctx since zq_create.  A run that replays a cached capture (same molecule and
 * options, no buffer of ctx re-allocated since) leaves it unchanged. */
int zq_scf_captures(const zq_ctx *ctx);

/* Diagnostics, with no counterpart in the reference: the SP2 densities of the last zq_scf_run_device on
 * ctx that the device-decided iteration finished (0s otherwise): how many, how many of them took
 * certified scale-and-fold steps (HOMO/LUMO bounds from an earlier plain density of the same SCF,
 * shifted by ||F' - F'_ref||_F), and the SP2 steps of all of them.  Returns ZQ_OK or ZQ_ERR_INVALID. */
int zq_scf_last_sp2(const zq_ctx *ctx, int *densities, int *folded, int *steps);

/* GPU symmetric eigensolver, eigh of src/linear_algebra.zig:45-71 (LAPACKE_dsyevd, row-major, 'V'):
 * A = U diag(W) U^T for the symmetric n x n A (row-major), W ascending, U (n x n row-major) with the
 * eigenvectors in its columns.  For n <= 192 the tridiagonal reduction is one workgroup holding A in
 * registers, the tridiagonal problem is rocSOLVER dstedc, and the back-transform one MFMA product;
 * larger n use rocSOLVER dsyevd.  zq_scf_run_device's final diagonalisation runs the same code on
 * device memory.  Host arrays; returns ZQ_OK, ZQ_ERR_LAPACK or a GPU error. */
int zq_eigh(zq_ctx *ctx, size_t n, const double *A, double *W, double *U);

/* Host generalized symmetric-definite eigensolver used by zq_scf_run (the dsygvd stand-in):
 * A x = w B x; on return A holds the eigenvectors (columns), w ascending; B is overwritten.
 * Returns 0, or ZQ_ERR_LAPACK. */
int zq_host_sygv(size_t n, double *A, double *B, double *w);

#ifdef __cplusplus
}
#endif

#endif /* ZINQ_SCF_H */
