/*
 * zinq_ints.h -- host Gaussian integral engine with the C ABI of the reference's libint shim.
 *
 * The reference reaches libint2 through src/libint.h (declared at src/libint.h:1-30, defined in
 * src/libint.cpp).  This library provides the same entry points without libint2, computed by an
 * in-house McMurchie-Davidson engine (s..g shells, pure spherical
 * d and higher, contracted shells renormalised).  The dense ERI layout is exactly the one
 * src/libint.cpp:143-150 writes: I[((p*n + q)*n + r)*n + s] = <pq|rs> = (pr|qs).
 *
 * Plain C99 so that Zig's @cImport can consume it.  Coordinates are in bohr; .xyz files are read in
 * Angstrom and converted with ZQ_BOHR_TO_ANGSTROM (CODATA 2018, the value libint2's read_dotxyz uses).
 */
#ifndef ZINQ_INTS_H
#define ZINQ_INTS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZQ_BOHR_TO_ANGSTROM 0.529177210903

typedef struct SystemData SystemData;

/* src/libint.cpp:10-18.  system: path of an .xyz file (Angstrom).  basis: path of a Gaussian94 basis
 * file, or "builtin:<name>" for the embedded sto-3g / cc-pvdz tables (the reference resolves
 * "builtin:" by exporting an embedded .g94 file, src/molecular_integrals.zig:96-125).  NULL on error. */
SystemData *libint_init(const char *system, const char *basis);
void libint_deinit(SystemData *sys);            /* src/libint.cpp:20-22 */
SystemData *libint_clone(SystemData *sys);      /* src/libint.cpp:24-26 */

size_t libint_nbf(SystemData *sys);             /* src/libint.cpp:32-34 */
size_t libint_nat(SystemData *sys);             /* src/libint.cpp:28-30 */

/* Dense outputs into caller-allocated buffers, exactly as src/libint.cpp:86-193. */
void libint_coulomb(double *I, SystemData *sys); /* n^4, physicist order, all 8 permutations */
void libint_kinetic(double *I, SystemData *sys); /* n^2 */
void libint_nuclear(double *I, SystemData *sys); /* n^2 */
void libint_overlap(double *I, SystemData *sys); /* n^2 */

void libint_atoms(int *atoms, SystemData *sys);       /* src/libint.cpp:36-42 */
void libint_coors(double *coors, SystemData *sys);    /* src/libint.cpp:44-52, bohr */
void libint_bf2at(int *bf2at, SystemData *sys);       /* src/libint.cpp:54-62 */
void libint_update_coords(SystemData *sys, const double *coords); /* src/libint.cpp:64-80 */

/* ---- extensions with no counterpart in src/libint.h ---- */

/* Build a system from atomic numbers and bohr coordinates (no file round trip). */
SystemData *zq_ints_from_atoms(size_t nat, const int *Z, const double *xyz_bohr, const char *basis);

/* Canonical 8-fold packed chemist-notation ERIs: out[I*(I+1)/2 + J] = (pq|rs) with p>=q, r>=s,
 * I = p*(p+1)/2 + q, J = r*(r+1)/2 + s, I >= J.  out has n(n+1)/2 * (n(n+1)/2 + 1) / 2 entries. */
void zq_ints_coulomb_packed(double *out, SystemData *sys);

/* Nuclear repulsion energy, src/molecular_system.zig:243-271. */
double zq_ints_nrep(SystemData *sys);

/* Number of electrons of the neutral system (sum of atomic numbers). */
size_t zq_ints_nel(SystemData *sys);

/* Host threads used by the ERI loops (default: $ZQ_NUM_THREADS, else $OMP_NUM_THREADS, else all). */
void zq_ints_set_threads(int nthreads);

/* Last error message of the calling thread (empty string if none). */
const char *zq_ints_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ZINQ_INTS_H */
