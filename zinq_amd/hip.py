"""ctypes bindings of libzinq_hip.so (include/zinq_hip.h and include/zinq_scf.h).

The calls are the GPU replacements of the reference's getFock (src/hartree_fock.zig:686-735),
ao2mo_pppp / ao2mo_oovv (src/integral_transform.zig:37-79,165-219) and the MP2 term of mp
(src/moller_plesset.zig:340-481).  Errors raise GpuError subclasses named after the Zig errors.
"""
import ctypes

import numpy as np

from ._native import load, ptr

ZQ_OK, ZQ_ERR_INVALID, ZQ_ERR_OOM, ZQ_ERR_LAUNCH, ZQ_ERR_NOT_CONVERGED, ZQ_ERR_LAPACK, ZQ_ERR_STATE = range(7)
BLOCK = {"full": 0, "ovov": 1}
PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2, "fp32c": 3}

EXPORTED = ["zq_create", "zq_destroy", "zq_last_error", "zq_eri_load", "zq_eri_load_packed", "zq_fock_jk", "zq_ao2mo",
            "zq_mp2", "zq_profile_enable", "zq_profile_reset", "zq_profile_get", "zq_device_bytes", "zq_scf_run",
            "zq_host_sygv", "zq_scf_run_device", "zq_scf_set_sp2_budget", "zq_eri_set_create", "zq_eri_set_destroy",
            "zq_eri_set_bytes", "zq_attach", "zq_mem_info", "zq_fock_jk_multi", "zq_eigh",
            "zq_scf_last_driver", "zq_scf_last_sp2", "zq_scf_captures", "zq_scf_set_sp2_guess_budget",
            "zq_scf_set_diis_error_scale", "zq_scf_last_diis_resets", "zq_scf_last_replays", "zq_mp2_noalloc",
            "zq_ao2mo_pp", "zq_ci", "zq_mp", "zq_grad_eri", "zq_ao2mo_dual", "zq_jk_record_count", "zq_mp2_resident"]


class GpuError(RuntimeError):
    code = None


class InvalidInput(GpuError):
    code = ZQ_ERR_INVALID


class GpuOutOfMemory(GpuError):
    code = ZQ_ERR_OOM


class GpuLaunch(GpuError):
    code = ZQ_ERR_LAUNCH


class ScfDidNotConverge(GpuError):
    code = ZQ_ERR_NOT_CONVERGED


class LapackError(GpuError):
    code = ZQ_ERR_LAPACK


class BadState(GpuError):
    code = ZQ_ERR_STATE


_ERRORS = {c.code: c for c in (InvalidInput, GpuOutOfMemory, GpuLaunch, ScfDidNotConverge, LapackError, BadState)}


class ScfOptions(ctypes.Structure):
    _fields_ = [("diis", ctypes.c_int), ("iterations", ctypes.c_int), ("threshold", ctypes.c_double),
                ("generalized", ctypes.c_int), ("fock_precision", ctypes.c_int)]


class ScfResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("iterations", ctypes.c_int)]


def lib():
    L = load("libzinq_hip.so")
    if not getattr(L, "_zq_typed", False):
        vp, sz, ci, cd = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
        L.zq_create.restype = vp
        L.zq_create.argtypes = [ci]
        L.zq_destroy.argtypes = [vp]
        L.zq_last_error.restype = ctypes.c_char_p
        L.zq_last_error.argtypes = [vp]
        L.zq_eri_load.argtypes = [vp, vp, sz]
        L.zq_eri_load_packed.argtypes = [vp, vp, sz]
        L.zq_fock_jk.argtypes = [vp, vp, vp, cd, vp]
        L.zq_ao2mo.argtypes = [vp, vp, sz, ci, ci, vp]
        L.zq_mp2.argtypes = [vp, vp, vp, sz, ci, vp]
        L.zq_mp2_noalloc.argtypes = [vp, vp, vp, sz, ci, vp]
        L.zq_mp2_resident.argtypes = [vp, sz, ci, vp]
        L.zq_ao2mo_pp.argtypes = [vp, sz, vp, vp, vp]
        L.zq_grad_eri.argtypes = [vp, vp, sz, sz, vp, cd, vp]
        L.zq_ao2mo_dual.argtypes = [vp, vp, vp, vp, sz, ci, vp, vp]
        L.zq_ci.argtypes = [vp, vp, vp, sz, ci, vp, sz, cd, sz, vp, vp]
        L.zq_mp.argtypes = [vp, vp, vp, sz, ci, ci, vp, vp]
        L.zq_profile_enable.argtypes = [vp, ci]
        L.zq_profile_reset.argtypes = [vp]
        L.zq_profile_get.argtypes = [vp, ctypes.c_char_p, vp, vp]
        L.zq_device_bytes.restype = sz
        L.zq_device_bytes.argtypes = [vp]
        L.zq_scf_run.argtypes = [vp, sz, sz, vp, vp, cd, ctypes.POINTER(ScfOptions), vp, vp, vp, vp,
                                 ctypes.POINTER(ScfResult)]
        L.zq_host_sygv.argtypes = [sz, vp, vp, vp]
        L.zq_scf_run_device.argtypes = L.zq_scf_run.argtypes
        L.zq_scf_set_sp2_budget.argtypes = [vp, ci]
        L.zq_eigh.argtypes = [vp, sz, vp, vp, vp]
        L.zq_scf_last_driver.argtypes = [vp]
        L.zq_scf_captures.argtypes = [vp]
        L.zq_scf_set_sp2_guess_budget.argtypes = [vp, ci]
        L.zq_scf_set_diis_error_scale.argtypes = [vp, cd]
        L.zq_scf_last_diis_resets.argtypes = [vp]
        L.zq_scf_last_replays.argtypes = [vp]
        L.zq_scf_last_sp2.argtypes = [vp, vp, vp, vp]
        L.zq_eri_set_create.restype = vp
        L.zq_eri_set_create.argtypes = [ci, vp, sz, ci]
        L.zq_eri_set_destroy.argtypes = [vp]
        L.zq_eri_set_bytes.restype = sz
        L.zq_eri_set_bytes.argtypes = [vp]
        L.zq_attach.argtypes = [vp, vp]
        L.zq_mem_info.argtypes = [ci, vp, vp]
        L.zq_jk_record_count.restype = sz
        L.zq_jk_record_count.argtypes = [sz]
        L.zq_fock_jk_multi.argtypes = [vp, vp, vp, sz, cd, vp]
        L._zq_typed = True
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def host_sygv(A, B):
    """The host dsygvd stand-in used by the SCF driver (A x = w B x)."""
    A, B = _f64(A).copy(), _f64(B).copy()
    w = np.zeros(A.shape[0])
    rc = lib().zq_host_sygv(A.shape[0], ptr(A), ptr(B), ptr(w))
    if rc:
        raise LapackError("generalized eigensolver failed")
    return w, A


def jk_record_count(n):
    """J/K records of the plan for n basis functions (zq_jk_record_count)."""
    return int(lib().zq_jk_record_count(n))


def mem_info(device=0):
    """(free, total) HBM bytes of the device (zq_mem_info)."""
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if lib().zq_mem_info(device, ctypes.byref(free), ctypes.byref(total)):
        raise GpuError("zq_mem_info failed")
    return free.value, total.value


class EriSet:
    """Resident ERIs of one molecule shared by any number of contexts (zq_eri_set_create)."""

    def __init__(self, packed, n, device=0, with_fp32=False):
        self._L = lib()
        pk = _f64(packed)
        h = self._L.zq_eri_set_create(device, ptr(pk), n, int(bool(with_fp32)))
        if not h:
            raise GpuOutOfMemory(f"zq_eri_set_create(n={n}) failed")
        self._h = ctypes.c_void_p(h)
        self.n = n
        self.nbytes = int(self._L.zq_eri_set_bytes(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self._L.zq_eri_set_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Context:
    """One HIP device context (own stream) holding resident ERIs: zq_create / zq_destroy."""

    def __init__(self, device=0):
        self._L = lib()
        h = self._L.zq_create(device)
        if not h:
            raise GpuError(f"zq_create({device}) failed: no usable HIP device")
        self._h = ctypes.c_void_p(h)
        self.n = None

    def close(self):
        if getattr(self, "_h", None):
            self._L.zq_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != ZQ_OK:
            msg = self._L.zq_last_error(self._h).decode()
            raise _ERRORS.get(rc, GpuError)(msg)

    def eri_load(self, g_dense):
        g = _f64(g_dense)
        n = g.shape[0]
        self._check(self._L.zq_eri_load(self._h, ptr(g), n))
        self.n = n

    def eri_load_packed(self, packed, n):
        pk = _f64(packed)
        if pk.size != (n * (n + 1) // 2) * (n * (n + 1) // 2 + 1) // 2:
            raise InvalidInput("packed ERI array has the wrong length")
        self._check(self._L.zq_eri_load_packed(self._h, ptr(pk), n))
        self.n = n

    def attach(self, eri_set):
        """Compute on a shared resident ERI set from now on (zq_attach)."""
        self._check(self._L.zq_attach(self._h, eri_set._h))
        self.n = eri_set.n

    def fock_jk(self, H, P, x_exch=0.5):
        H, P = _f64(H), _f64(P)
        F = np.zeros_like(H)
        self._check(self._L.zq_fock_jk(self._h, ptr(H), ptr(P), float(x_exch), ptr(F)))
        return F

    def fock_jk_multi(self, H, Ps, x_exch=0.5):
        """F[d] = H + G(P[d]) for a stack of densities in one ERI pass (H=None: bare G(P[d]))."""
        Ps = _f64(Ps)
        nd, n = Ps.shape[0], Ps.shape[1]
        F = np.zeros_like(Ps)
        Hp = ptr(_f64(H)) if H is not None else None
        self._check(self._L.zq_fock_jk_multi(self._h, Hp, ptr(Ps), nd, float(x_exch), ptr(F)))
        return F

    def ao2mo(self, C, nocc=0, block="full", precision="fp64"):
        C = _f64(C)
        n = C.shape[0]
        if block == "full":
            out = np.zeros((n, n, n, n))
        else:
            out = np.zeros((nocc, nocc, n - nocc, n - nocc))
        self._check(self._L.zq_ao2mo(self._h, ptr(C), nocc, BLOCK[block], PRECISION[precision], ptr(out)))
        return out

    def mp2(self, C, e, nocc, precision="fp64"):
        C, e = _f64(C), _f64(e)
        out = np.zeros(1)
        self._check(self._L.zq_mp2(self._h, ptr(C), ptr(e), nocc, PRECISION[precision], ptr(out)))
        return float(out[0])

    def mp2_resident(self, nocc, precision="fp64"):
        """MP2 correlation energy from the orbitals of the last device SCF on this context, which stay
        on the device (zq_mp2_resident); the same value as mp2() with the C and e it returned."""
        out = np.zeros(1)
        self._check(self._L.zq_mp2_resident(self._h, nocc, PRECISION[precision], ptr(out)))
        return float(out[0])

    def ao2mo_dual(self, dset, C, dC, nocc=0, block="full"):
        """(out, dout): the transform of g + eps dg (ctx's ERIs, dset's) with C + eps dC (zq_ao2mo_dual)."""
        C, dC = _f64(C), _f64(dC)
        n = C.shape[0]
        shape = (n, n, n, n) if block == "full" else (nocc, nocc, n - nocc, n - nocc)
        out, dout = np.zeros(shape), np.zeros(shape)
        self._check(self._L.zq_ao2mo_dual(self._h, dset._h, ptr(C), ptr(dC), nocc, BLOCK[block], ptr(out), ptr(dout)))
        return out, dout

    def grad_eri(self, dg_packed, n, P, x_exch=0.5):
        """sum 1/2 P P ((pq|rs)_x - x_exch (pr|qs)_x) for each packed derivative set dg_packed[x] (zq_grad_eri)."""
        dg, P = _f64(dg_packed), _f64(P)
        ncoord = dg.shape[0]
        out = np.zeros(ncoord)
        self._check(self._L.zq_grad_eri(self._h, ptr(dg), n, ncoord, ptr(P), float(x_exch), ptr(out)))
        return out

    def ao2mo_pp(self, A, C):
        """C^T A C on the GPU (zq_ao2mo_pp; ao2mo_pp of src/integral_transform.zig:12-34)."""
        A, C = _f64(A), _f64(C)
        out = np.zeros_like(A)
        self._check(self._L.zq_ao2mo_pp(self._h, A.shape[0], ptr(A), ptr(C), ptr(out)))
        return out

    def ci(self, H, C, nocc, excitations, VN, nstates=5, generalized=False):
        """Lowest CI state energies over the given excitation levels (zq_ci; computeCiStates of
        src/configuration_interaction.zig:597-683).  Returns (energies, number of determinants)."""
        H, C = _f64(H), _f64(C)
        exc = np.ascontiguousarray(excitations, dtype=np.int32)
        out = np.zeros(nstates)
        nd = ctypes.c_size_t()
        self._check(self._L.zq_ci(self._h, ptr(H), ptr(C), nocc, int(bool(generalized)), ptr(exc), len(exc), float(VN),
                                  nstates, ptr(out), ctypes.byref(nd)))
        return out[:min(nstates, nd.value)], nd.value

    def mp(self, C, e, nocc, order, generalized=False):
        """energies[0..order] of the determinant-space RS perturbation theory (zq_mp; mp of
        src/moller_plesset.zig:340-481).  Returns (energies, number of determinants)."""
        C, e = _f64(C), _f64(e)
        out = np.zeros(order + 1)
        nd = ctypes.c_size_t()
        self._check(self._L.zq_mp(self._h, ptr(C), ptr(e), nocc, int(bool(generalized)), int(order), ptr(out),
                                  ctypes.byref(nd)))
        return out, nd.value

    def mp2_noalloc(self, C, e, nocc, precision="fp64"):
        """zq_mp2 from the library's static scratch, allocating nothing (zq_mp2_noalloc; n <= 72)."""
        C, e = _f64(C), _f64(e)
        out = np.zeros(1)
        self._check(self._L.zq_mp2_noalloc(self._h, ptr(C), ptr(e), nocc, PRECISION[precision], ptr(out)))
        return float(out[0])

    def scf(self, H, S, VN, nocc, diis=8, iterations=100, threshold=1e-8, generalized=False, device=False,
            fock_precision="fp64"):
        """zq_scf_run (host DIIS/diagonalisation, reference-faithful) or, with device=True,
        zq_scf_run_device (everything n x n on the GPU; fock_precision="fp32" selects the mixed path)."""
        H, S = _f64(H), _f64(S)
        n = H.shape[0]
        C, e, P, F = np.zeros((n, n)), np.zeros(n), np.zeros((n, n)), np.zeros((n, n))
        opt = ScfOptions(int(diis or 0), int(iterations), float(threshold), int(bool(generalized)),
                         PRECISION[fock_precision])
        res = ScfResult()
        fn = self._L.zq_scf_run_device if device else self._L.zq_scf_run
        rc = fn(self._h, n, nocc, ptr(H), ptr(S), float(VN), ctypes.byref(opt), ptr(C), ptr(e), ptr(P),
                                ptr(F), ctypes.byref(res))
        self._check(rc)
        return dict(E=res.energy, iterations=res.iterations, C=C, e=e, P=P, F=F)

    def eigh(self, A):
        """W, U with A = U diag(W) U^T, W ascending (zq_eigh; eigh of src/linear_algebra.zig:45-71)."""
        A = _f64(A)
        n = A.shape[0]
        W, U = np.zeros(n), np.zeros((n, n))
        self._check(self._L.zq_eigh(self._h, n, ptr(A), ptr(W), ptr(U)))
        return W, U

    def last_scf_driver(self):
        """'device' or 'synced': the driver that finished the last device SCF (zq_scf_last_driver)."""
        return {0: None, 1: "device", 2: "synced"}[self._L.zq_scf_last_driver(self._h)]

    def scf_captures(self):
        """SCF iterations captured as hipGraphs by this context so far (zq_scf_captures)."""
        return int(self._L.zq_scf_captures(self._h))

    def last_scf_sp2(self):
        """{densities, folded, steps} of the last device SCF's SP2 runs (zq_scf_last_sp2)."""
        d, f, st = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.zq_scf_last_sp2(self._h, ctypes.byref(d), ctypes.byref(f), ctypes.byref(st)))
        return dict(densities=d.value, folded=f.value, steps=st.value)

    def scf_replays(self):
        """Iterations the device-decided driver enqueued in the last device SCF (zq_scf_last_replays)."""
        return int(self._L.zq_scf_last_replays(self._h))

    def set_sp2_guess_budget(self, steps):
        """SP2 steps for the device SCF's guess density (zq_scf_set_sp2_guess_budget)."""
        self._check(self._L.zq_scf_set_sp2_guess_budget(self._h, int(steps)))

    def set_diis_error_scale(self, scale):
        """Factor on the device SCF's stored DIIS error vectors, in [0, 1] (zq_scf_set_diis_error_scale);
        0 makes every DIIS system singular, so each iteration takes the history reset."""
        self._check(self._L.zq_scf_set_diis_error_scale(self._h, float(scale)))

    def last_diis_resets(self):
        """Singular DIIS systems the last device SCF met (zq_scf_last_diis_resets)."""
        return int(self._L.zq_scf_last_diis_resets(self._h))

    def set_sp2_budget(self, steps):
        """SP2 steps per density in the captured device-SCF iteration (zq_scf_set_sp2_budget)."""
        self._check(self._L.zq_scf_set_sp2_budget(self._h, int(steps)))

    def profile(self, on=True):
        self._L.zq_profile_enable(self._h, int(on))

    def profile_reset(self):
        self._L.zq_profile_reset(self._h)

    def profile_get(self, prefix):
        ms, cnt = ctypes.c_double(), ctypes.c_int()
        self._L.zq_profile_get(self._h, prefix.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        return ms.value, cnt.value

    def device_bytes(self):
        return int(self._L.zq_device_bytes(self._h))
