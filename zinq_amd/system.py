"""MolecularSystem: the host-side producer of the hot path's inputs.

Mirrors src/molecular_system.zig (init :27-80, coulomb :115-122, coulombSpin :161-182, nrep :243-271)
on top of libzinq_ints.so, whose C ABI is the reference's src/libint.h.  Integrals are FP64, row-major:
``coulomb()`` is the dense physicist tensor g[p,q,r,s] = <pq|rs> = (pr|qs) that src/libint.cpp:143-150
writes; ``coulomb_packed()`` is the canonical 8-fold packed chemist form the GPU path loads.
"""
import ctypes

import numpy as np

from ._native import load, ptr


def _lib():
    lib = load("libzinq_ints.so")
    if not getattr(lib, "_zq_typed", False):
        lib.libint_init.restype = ctypes.c_void_p
        lib.libint_init.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        lib.zq_ints_from_atoms.restype = ctypes.c_void_p
        lib.zq_ints_from_atoms.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p]
        lib.libint_clone.restype = ctypes.c_void_p
        lib.libint_clone.argtypes = [ctypes.c_void_p]
        lib.libint_deinit.argtypes = [ctypes.c_void_p]
        for f in ("libint_nbf", "libint_nat", "zq_ints_nel"):
            getattr(lib, f).restype = ctypes.c_size_t
            getattr(lib, f).argtypes = [ctypes.c_void_p]
        for f in ("libint_coulomb", "libint_kinetic", "libint_nuclear", "libint_overlap", "zq_ints_coulomb_packed",
                  "libint_atoms", "libint_coors", "libint_bf2at"):
            getattr(lib, f).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            getattr(lib, f).restype = None
        lib.libint_update_coords.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        lib.zq_ints_nrep.restype = ctypes.c_double
        lib.zq_ints_nrep.argtypes = [ctypes.c_void_p]
        lib.zq_ints_set_threads.argtypes = [ctypes.c_int]
        lib.zq_ints_last_error.restype = ctypes.c_char_p
        lib._zq_typed = True
    return lib


class InitializationFailed(RuntimeError):
    """error.InitializationFailed of src/molecular_system.zig:34."""


class InvalidInput(ValueError):
    """error.InvalidInput of src/molecular_system.zig:57-75."""


def npair(n):
    return n * (n + 1) // 2


def npacked(n):
    p = npair(n)
    return p * (p + 1) // 2


class MolecularSystem:
    def __init__(self, handle, charge=0, multiplicity=1):
        self._lib = _lib()
        self._h = ctypes.c_void_p(handle)
        nat = self._lib.libint_nat(self._h)
        self.nbf = int(self._lib.libint_nbf(self._h))
        self.atoms = np.zeros(nat, dtype=np.int32)
        self.coors = np.zeros(3 * nat)
        self._lib.libint_atoms(ptr(self.atoms), self._h)
        self._lib.libint_coors(ptr(self.coors), self._h)
        nel = int(self.atoms.sum()) - charge
        if nel <= 0:
            raise InvalidInput("NUMBER OF ELECTRONS MUST BE GREATER THAN 0")
        if (nel + multiplicity - 1) % 2 != 0:
            raise InvalidInput("CHARGE AND MULTIPLICITY ARE INCOMPATIBLE WITH THE NUMBER OF ELECTRONS")
        if nel + 1 < multiplicity:
            raise InvalidInput("MULTIPLICITY IS INCOMPATIBLE WITH THE NUMBER OF ELECTRONS")
        self.nel = nel

    @classmethod
    def from_xyz(cls, system, basis, charge=0, multiplicity=1):
        """src/molecular_system.zig:27 init(system, basis, charge, multiplicity)."""
        h = _lib().libint_init(system.encode(), basis.encode())
        if not h:
            raise InitializationFailed(_lib().zq_ints_last_error().decode())
        return cls(h, charge, multiplicity)

    @classmethod
    def from_atoms(cls, Z, xyz_bohr, basis, charge=0, multiplicity=1):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        h = _lib().zq_ints_from_atoms(len(Z), ptr(Z), ptr(xyz), basis.encode())
        if not h:
            raise InitializationFailed(_lib().zq_ints_last_error().decode())
        return cls(h, charge, multiplicity)

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.libint_deinit(self._h)
            self._h = None

    def _one(self, fn):
        out = np.zeros((self.nbf, self.nbf))
        getattr(self._lib, fn)(ptr(out), self._h)
        return out

    def overlap(self):
        return self._one("libint_overlap")

    def kinetic(self):
        return self._one("libint_kinetic")

    def nuclear(self):
        return self._one("libint_nuclear")

    def hcore(self):
        """H = K + V, src/molecular_integrals.zig:198-204."""
        return self.kinetic() + self.nuclear()

    def coulomb(self):
        n = self.nbf
        out = np.zeros((n, n, n, n))
        self._lib.libint_coulomb(ptr(out), self._h)
        return out

    def coulomb_packed(self):
        out = np.zeros(npacked(self.nbf))
        self._lib.zq_ints_coulomb_packed(ptr(out), self._h)
        return out

    def nrep(self):
        return float(self._lib.zq_ints_nrep(self._h))


def spin_block_1e(A):
    """overlapSpin/kineticSpin/nuclearSpin of src/molecular_system.zig: diag(A, A)."""
    n = A.shape[0]
    out = np.zeros((2 * n, 2 * n))
    out[:n, :n] = A
    out[n:, n:] = A
    return out


def spin_block_2e(g):
    """coulombSpin, src/molecular_system.zig:161-182 (physicist order)."""
    n = g.shape[0]
    out = np.zeros((2 * n,) * 4)
    out[:n, :n, :n, :n] = g
    out[n:, :n, n:, :n] = g
    out[:n, n:, :n, n:] = g
    out[n:, n:, n:, n:] = g
    return out
