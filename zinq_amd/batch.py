"""ctypes bindings of the grouped many-molecule RHF + MP2 (include/zinq_batch.h)."""
import ctypes

import numpy as np

from ._native import ptr
from .hip import PRECISION, ScfOptions, GpuError, _ERRORS, _f64, lib as _hip_lib

EXPORTED = ["zq_batch_create", "zq_batch_destroy", "zq_batch_last_error", "zq_batch_add", "zq_batch_count",
            "zq_batch_autotune", "zq_batch_run", "zq_batch_orbitals", "zq_batch_timings"]


def lib():
    L = _hip_lib()
    if not getattr(L, "_zq_batch_typed", False):
        vp, sz, ci, cd = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
        L.zq_batch_create.restype = vp
        L.zq_batch_create.argtypes = [ci]
        L.zq_batch_destroy.argtypes = [vp]
        L.zq_batch_last_error.restype = ctypes.c_char_p
        L.zq_batch_last_error.argtypes = [vp]
        L.zq_batch_add.argtypes = [vp, sz, sz, vp, vp, cd, vp]
        L.zq_batch_count.argtypes = [vp]
        L.zq_batch_autotune.argtypes = [vp, vp]
        L.zq_batch_run.argtypes = [vp, ctypes.POINTER(ScfOptions), ci, vp, vp, vp, vp]
        L.zq_batch_orbitals.argtypes = [vp, ci, vp, vp]
        L.zq_batch_timings.argtypes = [vp, vp]
        L._zq_batch_typed = True
    return L


class Batch:
    """Many closed-shell molecules (n <= 96) through one grouped, graph-captured RHF + MP2."""

    def __init__(self, device=0):
        self._L = lib()
        h = self._L.zq_batch_create(device)
        if not h:
            raise GpuError(f"zq_batch_create({device}) failed: no usable HIP device")
        self._h = ctypes.c_void_p(h)
        self.sizes = []

    def close(self):
        if getattr(self, "_h", None):
            self._L.zq_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _err(self, rc):
        raise _ERRORS.get(rc, GpuError)(self._L.zq_batch_last_error(self._h).decode())

    def add(self, H, S, VN, nocc, eri_packed):
        H, S, pk = _f64(H), _f64(S), _f64(eri_packed)
        n = H.shape[0]
        rc = self._L.zq_batch_add(self._h, n, nocc, ptr(H), ptr(S), float(VN), ptr(pk))
        if rc < 0:
            self._err(-rc)
        self.sizes.append(n)
        return rc

    def add_system(self, system):
        return self.add(system.hcore(), system.overlap(), system.nrep(), system.nel // 2, system.coulomb_packed())

    def autotune(self):
        chosen = np.zeros(2, dtype=np.int32)
        rc = self._L.zq_batch_autotune(self._h, ptr(chosen))
        if rc:
            self._err(rc)
        return dict(item_quartets=int(chosen[0]), workgroup=int(chosen[1]))

    def run(self, diis=8, iterations=100, threshold=1e-8, mp2_precision="fp64"):
        m = len(self.sizes)
        E, Ec = np.zeros(m), np.zeros(m)
        it, st = np.zeros(m, dtype=np.int32), np.zeros(m, dtype=np.int32)
        opt = ScfOptions(int(diis or 0), int(iterations), float(threshold), 0, 0)
        rc = self._L.zq_batch_run(self._h, ctypes.byref(opt), PRECISION[mp2_precision], ptr(E), ptr(Ec), ptr(it), ptr(st))
        if rc:
            self._err(rc)
        return dict(E_rhf=E, E_corr=Ec, iterations=it, status=st)

    def timings(self):
        """Stage times of the last run, ms (host wall clock): scf, eigen, mp2, total (zq_batch_timings)."""
        ms = np.zeros(4)
        rc = self._L.zq_batch_timings(self._h, ptr(ms))
        if rc:
            self._err(rc)
        return dict(zip(("scf", "eigen", "mp2", "total"), (round(float(x), 3) for x in ms)))

    def orbitals(self, index):
        n = self.sizes[index]
        e, C = np.zeros(n), np.zeros((n, n))
        rc = self._L.zq_batch_orbitals(self._h, int(index), ptr(e), ptr(C))
        if rc:
            self._err(rc)
        return e, C
