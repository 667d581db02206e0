/*
 * zinq_batch.h -- many small molecules through one grouped, graph-captured RHF + MP2.
 *
 * The reference runs one molecule per process through scf (src/hartree_fock.zig:738-875), getFock
 * (:686-735) and mp (src/moller_plesset.zig:340-481).  A batch runs the same arithmetic (FP64, the
 * reference's DIIS and convergence test) for many molecules of uneven size at once:
 *   - the ERIs of every molecule stay resident in one buffer of 8x8 AO-block quartets;
 *   - one J/K launch (records + reduce) serves every molecule still iterating, from one grouped work
 *     plan whose item size is chosen by zq_batch_autotune;
 *   - one launch runs the rest of the iteration for every molecule, one workgroup per molecule
 *     (energy, DIIS error/history/LU/extrapolation, X^T F X, SP2 density, convergence test);
 *   - the iteration is captured as one hipGraph and replayed until every molecule has converged;
 *   - the final eigenpairs come from batched rocSOLVER dsyevd per distinct size; the MP2 energies
 *     from the OVOV transform of zinq_hip.h grouped over all molecules: one launch per stage (C slices,
 *     Q1..Q4, pair sums, totals), with C and the orbital energies kept on the device.  Each molecule's
 *     energy is bitwise the one zq_mp2 gives for its orbitals (FP64, FP32, FP32C; BF16X3 runs the
 *     single-molecule transform for several molecules at a time on their own streams).
 * Limits: RHF, n <= 96 basis functions per molecule, DIIS history <= 8.
 */
#ifndef ZINQ_BATCH_H
#define ZINQ_BATCH_H

#include <stddef.h>

#include "zinq_scf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct zq_batch zq_batch;

/* Batch on HIP device `device`; NULL on failure. */
zq_batch *zq_batch_create(int device);
void zq_batch_destroy(zq_batch *b);
const char *zq_batch_last_error(const zq_batch *b);

/* Adds one closed-shell molecule: n basis functions (n <= 96), nocc doubly occupied orbitals, H and S
 * (n x n, row-major), V_NN, and its canonical packed ERIs (the zq_eri_load_packed layout).  The inputs
 * are copied.  Returns the molecule's index (>= 0) or a negative ZQ_ERR_* code. */
int zq_batch_add(zq_batch *b, size_t n, size_t nocc, const double *H, const double *S, double VN,
                 const double *eri_packed);
int zq_batch_count(const zq_batch *b);

/* Times the grouped J/K build for each candidate item size (quartets per work item) and the per-molecule
 * step for each workgroup size on the resident batch, and keeps the fastest.  Optional; uploads the
 * batch if needed.  chosen (may be NULL) receives {item size, workgroup size}. */
int zq_batch_autotune(zq_batch *b, int chosen[2]);

/* RHF (opt: diis <= 8, iterations, threshold; generalized and fock_precision must be 0) and the MP2
 * correlation energy (transform precision as in zq_ao2mo) of every molecule.  Arrays of length
 * zq_batch_count: E_rhf (total RHF energy), E_corr, iterations, status (ZQ_OK or ZQ_ERR_NOT_CONVERGED /
 * ZQ_ERR_LAPACK for that molecule).  Returns ZQ_OK when the batch ran (per-molecule status aside). */
int zq_batch_run(zq_batch *b, const zq_scf_options *opt, int mp2_precision, double *E_rhf, double *E_corr,
                 int *iterations, int *status);

/* Stage times of the last zq_batch_run, host wall clock in ms: {SCF iterations, final eigenpairs and
 * orbitals, MP2, total}.  Diagnostics, with no counterpart in the reference. */
int zq_batch_timings(const zq_batch *b, double ms[4]);

/* Orbital energies and coefficients of molecule `index` from the last zq_batch_run: e (n), C (n x n). */
int zq_batch_orbitals(const zq_batch *b, int index, double *e, double *C);

#ifdef __cplusplus
}
#endif

#endif /* ZINQ_BATCH_H */
