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

/* Same contract as zq_scf_run with every n x n step on the GPU (J/K, energy, DIIS error and
 * extrapolation, density); the diagonalisation is a rocSOLVER dsyevd of F in the previous MO basis
 * (C_prev^T F C_prev, with C_prev^T S C_prev = I), C = C_prev C'.  Only the <= 9 x 9
 * DIIS system and the convergence test run on the host. */
int zq_scf_run_device(zq_ctx *ctx, size_t n, size_t nocc, const double *H, const double *S, double VN,
                      const zq_scf_options *opt, double *C, double *e, double *P, double *F, zq_scf_result *res);

/* Host generalized symmetric-definite eigensolver used by zq_scf_run (the dsygvd stand-in):
 * A x = w B x; on return A holds the eigenvectors (columns), w ascending; B is overwritten.
 * Returns 0, or ZQ_ERR_LAPACK. */
int zq_host_sygv(size_t n, double *A, double *B, double *w);

#ifdef __cplusplus
}
#endif

#endif /* ZINQ_SCF_H */
