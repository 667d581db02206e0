"""ctypes bindings of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by zinq_amd.  See oracle.c for the reference file:line each function restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

c_int, c_double, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p


def build():
    """Compile liboracle.so with the committed Makefile (gcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _lib = ctypes.CDLL(path)
        L = _lib
        L.orc_get_fock.argtypes = [c_int, vp, vp, vp, c_int, vp]
        L.orc_get_energy.argtypes = [c_int, vp, vp, vp]
        L.orc_get_energy.restype = c_double
        L.orc_get_density.argtypes = [c_int, vp, c_int, c_int, vp]
        L.orc_get_density.restype = c_double
        L.orc_get_error.argtypes = [c_int, vp, vp, vp, vp]
        L.orc_diis.argtypes = [c_int, c_int, vp, vp, vp]
        L.orc_geigh.argtypes = [c_int, vp, vp, vp, vp]
        L.orc_eigh.argtypes = [c_int, vp, vp, vp]
        L.orc_scf.argtypes = [c_int, c_int, c_int, vp, vp, vp, c_double, c_int, c_int, c_double, vp, vp, vp, vp, vp, vp]
        L.orc_ao2mo_pp.argtypes = [c_int, vp, vp, vp]
        L.orc_ao2mo_pppp.argtypes = [c_int, vp, vp, vp]
        L.orc_ao2mo_oovv.argtypes = [c_int, c_int, vp, vp, vp]
        L.orc_ao2so_coef.argtypes = [c_int, vp, vp]
        L.orc_ao2so_pp.argtypes = [c_int, vp, vp]
        L.orc_ao2so_pppp.argtypes = [c_int, vp, vp]
        L.orc_mp.argtypes = [c_int, c_int, vp, vp, vp, c_int, c_int, vp]
        L.orc_mp2_closed.argtypes = [c_int, c_int, vp, vp]
        L.orc_mp2_closed.restype = c_double
        L.orc_ci.argtypes = [c_int, vp, vp, vp, c_int, c_int, vp, c_int, c_double, c_int, vp]
        L.orc_get_fock_par.argtypes = [c_int, vp, vp, vp, c_int, vp]
        L.orc_ao2mo_oovv_par.argtypes = [c_int, c_int, vp, vp, vp]
        L.orc_max_threads.restype = c_int
        L.orc_fock_rows_packed.argtypes = [c_int, vp, vp, vp, c_double, c_int, vp, vp]
        L.orc_ovov_pairs_packed.argtypes = [c_int, c_int, vp, vp, c_int, vp, vp, vp]
        L.orc_get_fock_packed.argtypes = [c_int, vp, vp, vp, c_double, vp]
        L.orc_scf_packed.argtypes = L.orc_scf.argtypes
        L.orc_scf_par.argtypes = L.orc_scf.argtypes
        L.orc_ao2mo_pppp_dual.argtypes = [c_int, vp, vp, vp, vp, vp, vp]
        L.orc_grad_eri.argtypes = [c_int, c_int, vp, vp, c_int, c_double, vp]
        L.orc_get_fock_dft.argtypes = [c_int, vp, vp, vp, c_int, c_double, vp, vp]
    return _lib


def max_threads():
    """Threads the *_par functions use (OpenMP)."""
    return int(lib().orc_max_threads())


def get_fock_par(g, H, P, generalized=False):
    """get_fock on all cores; bitwise equal to get_fock (same per-element accumulation order)."""
    g, H, P = _c(g), _c(H), _c(P)
    n = H.shape[0]
    F = np.zeros((n, n))
    lib().orc_get_fock_par(n, _p(g), _p(H), _p(P), int(generalized), _p(F))
    return F


def ao2mo_oovv_par(g, C, nocc):
    """ao2mo_oovv on all cores; bitwise equal to ao2mo_oovv."""
    g, C = _c(g), _c(C)
    n = C.shape[0]
    out = np.zeros((nocc, nocc, n - nocc, n - nocc))
    lib().orc_ao2mo_oovv_par(n, nocc, _p(g), _p(C), _p(out))
    return out


def _p(a):
    assert a.flags["C_CONTIGUOUS"] and a.dtype in (np.float64, np.int32)
    return ctypes.c_void_p(a.ctypes.data)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def get_fock(g, H, P, generalized=False):
    n = H.shape[0]
    F = np.zeros((n, n))
    lib().orc_get_fock(n, _p(_c(g)), _p(_c(H)), _p(_c(P)), int(generalized), _p(F))
    return F


def get_energy(H, F, P):
    return lib().orc_get_energy(H.shape[0], _p(_c(H)), _p(_c(F)), _p(_c(P)))


def get_density(C, nocc, generalized=False, P_old=None):
    n = C.shape[0]
    P = np.zeros((n, n)) if P_old is None else _c(P_old).copy()
    rms = lib().orc_get_density(n, _p(_c(C)), nocc, int(generalized), _p(P))
    return P, rms


def get_error(F, P, S):
    n = F.shape[0]
    e = np.zeros((n, n))
    lib().orc_get_error(n, _p(_c(F)), _p(_c(P)), _p(_c(S)), _p(e))
    return e


def diis(fhist, ehist):
    """Returns (F, ok); ok is False when the LU solve fails (src/hartree_fock.zig:821)."""
    m, n = len(fhist), fhist[0].shape[0]
    fh, eh = _c(np.stack(fhist)), _c(np.stack(ehist))
    F = np.zeros((n, n))
    rc = lib().orc_diis(m, n, _p(fh), _p(eh), _p(F))
    return F, rc == 0


def geigh(A, B):
    n = A.shape[0]
    e, C = np.zeros(n), np.zeros((n, n))
    rc = lib().orc_geigh(n, _p(_c(A)), _p(_c(B)), _p(e), _p(C))
    if rc:
        raise RuntimeError("LapackError")
    return e, C


def eigh(A):
    n = A.shape[0]
    e, V = np.zeros(n), np.zeros((n, n))
    lib().orc_eigh(n, _p(_c(A).copy()), _p(e), _p(V))
    return e, V


class ScfDidNotConverge(RuntimeError):
    pass


def scf(H, S, g, VN, nocc, generalized=False, diis=8, iterations=100, threshold=1e-8, packed=None, par=False):
    """src/hartree_fock.zig:738-875 with the core guess of :425-431.  Returns dict(E, C, e, P, F, iterations).
    With packed= (canonical packed ERIs, g ignored) the Fock builds read those; with par=True they run on
    all cores.  Both are bitwise the faithful single-core run."""
    n = H.shape[0]
    fn = lib().orc_scf_par if par else lib().orc_scf
    if packed is not None:
        g, fn = packed, lib().orc_scf_packed
    E = np.zeros(1)
    C, e, P, F = np.zeros((n, n)), np.zeros(n), np.zeros((n, n)), np.zeros((n, n))
    it = np.zeros(1, dtype=np.int32)
    rc = fn(n, nocc, int(generalized), _p(_c(H)), _p(_c(S)), _p(_c(g)), VN, diis or 0, iterations, threshold,
                       _p(E), _p(C), _p(e), _p(P), _p(F), _p(it))
    if rc == 1:
        raise ScfDidNotConverge()
    if rc == 2:
        raise RuntimeError("LapackError")
    return dict(E=float(E[0]), C=C, e=e, P=P, F=F, iterations=int(it[0]))


def get_fock_packed(packed, H, P, x_exch=0.5):
    """getFock from canonical packed ERIs; bitwise get_fock."""
    n = H.shape[0]
    F = np.zeros((n, n))
    lib().orc_get_fock_packed(n, _p(_c(packed)), _p(_c(H)), _p(_c(P)), float(x_exch), _p(F))
    return F


def fock_rows_packed(packed, H, P, rows, x_exch=0.5):
    """Rows of getFock's symmetrised F read from canonical packed ERIs; bitwise those of get_fock."""
    n = H.shape[0]
    r = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.zeros((len(r), n))
    lib().orc_fock_rows_packed(n, _p(_c(packed)), _p(_c(H)), _p(_c(P)), float(x_exch), len(r), _p(r), _p(out))
    return out


def ovov_pairs_packed(packed, C, nocc, pairs):
    """(ia|jb) blocks [t][a][b] of ao2mo_oovv for occupied pairs (i, j) from canonical packed ERIs; bitwise
    those of ao2mo_oovv.  Pairs with the same i should be adjacent (the first quarter is kept per i)."""
    n = C.shape[0]
    pi = np.ascontiguousarray([p[0] for p in pairs], dtype=np.int32)
    pj = np.ascontiguousarray([p[1] for p in pairs], dtype=np.int32)
    out = np.zeros((len(pairs), n - nocc, n - nocc))
    lib().orc_ovov_pairs_packed(n, nocc, _p(_c(packed)), _p(_c(C)), len(pairs), _p(pi), _p(pj), _p(out))
    return out


def grad_eri(dg, P, generalized=False, exx=1.0):
    """Two-electron part of the HF gradient for a stack of dense derivative tensors dg[x]."""
    dg = _c(dg)
    ncoord, n = dg.shape[0], dg.shape[1]
    out = np.zeros(ncoord)
    lib().orc_grad_eri(n, ncoord, _p(dg), _p(_c(P)), int(generalized), float(exx), _p(out))
    return out


def get_fock_dft(g, H, P, exx, Vxc=None, generalized=False):
    """getFock's DFT branch (Coulomb loop, exx-scaled exchange loop, + Vxc, symmetrised)."""
    n = H.shape[0]
    F = np.zeros((n, n))
    V = _p(_c(Vxc)) if Vxc is not None else None
    lib().orc_get_fock_dft(n, _p(_c(g)), _p(_c(H)), _p(_c(P)), int(generalized), float(exx), V, _p(F))
    return F


def ao2mo_pp(A, C):
    N = A.shape[0]
    out = np.zeros((N, N))
    lib().orc_ao2mo_pp(N, _p(_c(A)), _p(_c(C)), _p(out))
    return out


def ao2mo_pppp(g, C):
    N = C.shape[0]
    out = np.zeros((N,) * 4)
    lib().orc_ao2mo_pppp(N, _p(_c(g)), _p(_c(C)), _p(out))
    return out


def ao2mo_pppp_dual(g, dg, C, dC):
    """ao2mo_pppp over dual numbers: (value, derivative) of the transform of g + eps dg with C + eps dC."""
    N = C.shape[0]
    out, dout = np.zeros((N,) * 4), np.zeros((N,) * 4)
    lib().orc_ao2mo_pppp_dual(N, _p(_c(g)), _p(_c(dg)), _p(_c(C)), _p(_c(dC)), _p(out), _p(dout))
    return out, dout


def ao2mo_oovv(g, C, nocc):
    N = C.shape[0]
    out = np.zeros((nocc, nocc, N - nocc, N - nocc))
    lib().orc_ao2mo_oovv(N, nocc, _p(_c(g)), _p(_c(C)), _p(out))
    return out


def ao2so_coef(C):
    n = C.shape[0]
    out = np.zeros((2 * n, 2 * n))
    lib().orc_ao2so_coef(n, _p(_c(C)), _p(out))
    return out


def mp(order, g, C, e, nocc, generalized=False):
    """src/moller_plesset.zig:340-481; returns energies[0..order]."""
    n = C.shape[0]
    out = np.zeros(order + 1)
    lib().orc_mp(order, n, _p(_c(g)), _p(_c(C)), _p(_c(e)), nocc, int(generalized), _p(out))
    return out


def mp2_closed(oovv, e):
    o, _, v, _ = oovv.shape
    return lib().orc_mp2_closed(o, v, _p(_c(oovv)), _p(_c(e)))


def ci(g, H, C, nocc, excitations, VN, nstates=5, generalized=False):
    """src/configuration_interaction.zig:597-683; lowest nstates total energies."""
    n = C.shape[0]
    exc = np.ascontiguousarray(excitations, dtype=np.int32)
    out = np.zeros(nstates)
    lib().orc_ci(n, _p(_c(g)), _p(_c(H)), _p(_c(C)), nocc, int(generalized), _p(exc), len(exc), VN, nstates, _p(out))
    return out
