"""Serving loop: RHF + MP2 requests with Poisson arrivals, continuous batching on one GPU.

A request names a molecule (host H, S, V_NN and packed ERIs, generated before the run).  The server
keeps a pool of `slots` contexts, each with its own HIP stream, and an LRU cache of resident ERI sets
(zq_eri_set) shared by the contexts.  A request is admitted when a slot is free and its footprint fits
the HBM budget: the molecule's resident set (8 N_u bytes of quartets, plus 4 N_u for the FP32 copy of
the mixed path) unless already resident.  The budget is 85% of the free HBM less, per slot, the
transient scratch of the largest request (J/K records and the OVOV transform intermediates), which a
slot's context keeps once grown.  Sets not used by a running request are evicted oldest-first.
Admitted requests run concurrently on their slots' streams, so their kernels overlap on the device;
a request that arrives while others run starts as soon as a slot and memory free up (continuous
batching, not a queue drained in rounds).  Latency = completion - arrival, per request.
"""
import collections
import math
import threading
import time

import numpy as np

from . import hip


def eri_set_bytes(n, with_fp32):
    nb = (n + 7) // 8
    Nb = nb * (nb + 1) // 2
    nq = Nb * (Nb + 1) // 2
    return nq * 4096 * (8 + (4 if with_fp32 else 0))


def scratch_bytes(n, o, transform_bytes=4):
    """Transient device memory of one request: the J/K records of the molecule's plan (64 doubles each)
    and the OVOV intermediates X, Y, W, OUT of the transform (element size transform_bytes)."""
    nb = (n + 7) // 8
    npd, Nb = nb * 8, nb * (nb + 1) // 2
    v = n - o
    records = hip.jk_record_count(n) * 64 * 8
    tr = (Nb * 64 * o * npd + o * o * npd * npd + o * o * npd * v + o * o * v * v) * transform_bytes
    return records + tr + 64 * n * n * 8


class Molecule:
    def __init__(self, key, system):
        self.key = key
        self.n, self.o = system.nbf, system.nel // 2
        self.H, self.S, self.VN = system.hcore(), system.overlap(), system.nrep()
        self.packed = system.coulomb_packed()


class Server:
    def __init__(self, molecules, slots=8, device=0, hbm_budget=None, mode=None):
        self.mols = {m.key: m for m in molecules}
        self.mode = mode or dict(fock="fp32", transform="fp32c", threshold=1e-6)
        # the FP32 ERI copy lives in the shared set whenever a Fock build or the transform reads it, so no
        # slot context makes (and frees) a copy of its own
        self.with_fp32 = self.mode["fock"] == "fp32" or self.mode["transform"] in ("fp32", "fp32c")
        self.tbytes = 8 if self.mode["transform"] == "fp64" else 4
        if hbm_budget is None:
            hbm_budget = int(0.85 * hip.mem_info(device)[0])
        # every slot's context keeps the scratch of the largest request it served: reserve that up front
        self.slot_reserve = max(scratch_bytes(m.n, m.o, self.tbytes) for m in molecules)
        self.budget = hbm_budget - slots * self.slot_reserve
        if self.budget <= 0:
            raise hip.GpuOutOfMemory("HBM budget below the slots' scratch reserve")
        self.ctx = [hip.Context(device) for _ in range(slots)]
        self.device = device
        self.free_slots = list(range(slots))
        self.resident = collections.OrderedDict()  # key -> EriSet (LRU order)
        self.in_use = collections.Counter()          # key -> running requests
        self.reserved = 0                            # bytes of resident sets
        self.lock = threading.Condition()

    def close(self):
        for c in self.ctx:
            c.close()
        for s in self.resident.values():
            s.close()

    # ---- admission (called with the lock held)
    def _try_admit(self, key):
        if not self.free_slots:
            return None
        m = self.mols[key]
        if key not in self.resident:
            need = eri_set_bytes(m.n, self.with_fp32)
            # evict idle sets, oldest first, until the molecule's set fits
            while self.reserved + need > self.budget:
                idle = [k for k in self.resident if self.in_use[k] == 0]
                if not idle:
                    return None
                victim = idle[0]
                self.resident.pop(victim).close()
                self.reserved -= eri_set_bytes(self.mols[victim].n, self.with_fp32)
            self.resident[key] = hip.EriSet(m.packed, m.n, self.device, self.with_fp32)
            self.reserved += eri_set_bytes(m.n, self.with_fp32)
        self.resident.move_to_end(key)
        self.in_use[key] += 1
        return self.free_slots.pop()

    def _release(self, slot, key):
        self.in_use[key] -= 1
        self.free_slots.append(slot)
        self.lock.notify_all()

    def _run(self, slot, key):
        m, c, md = self.mols[key], self.ctx[slot], self.mode
        c.attach(self.resident[key])
        r = c.scf(m.H, m.S, m.VN, m.o, diis=8, iterations=100, threshold=md["threshold"], device=True,
                  fock_precision=md["fock"])
        e2 = c.mp2_resident(m.o, md["transform"])  # the SCF's orbitals, still on the device
        return dict(E_rhf=r["E"], E_corr=e2, E_mp2=r["E"] + e2, iterations=r["iterations"])

    def serve(self, requests):
        """requests: list of (arrival_s, key) sorted by arrival.  Returns per-request records."""
        out = [None] * len(requests)
        pending = collections.deque(range(len(requests)))
        t0 = time.perf_counter()
        workers = []

        def work(i, slot):
            key = requests[i][1]
            start = time.perf_counter() - t0
            try:
                res = self._run(slot, key)
            except Exception as e:  # recorded, never silent
                res = dict(error=repr(e))
            done = time.perf_counter() - t0
            res.update(key=key, arrival=requests[i][0], start=start, done=done, latency=done - requests[i][0],
                       slot=slot)
            out[i] = res
            with self.lock:
                self._release(slot, key)

        with self.lock:
            while pending:
                i = pending[0]
                now = time.perf_counter() - t0
                wait = requests[i][0] - now
                if wait > 0:
                    self.lock.wait(timeout=wait)
                    continue
                key = requests[i][1]
                fail = None
                if key not in self.mols:
                    fail = KeyError(f"unknown molecule {key!r}")
                elif key not in self.resident and eri_set_bytes(self.mols[key].n, self.with_fp32) > self.budget:
                    # can never be admitted, even with every other set evicted
                    fail = hip.GpuOutOfMemory(f"{key}: ERI set above the HBM budget")
                else:
                    try:
                        slot = self._try_admit(key)
                    except Exception as e:  # the set could not be created: recorded, never silent
                        fail = e
                if fail is not None:
                    out[i] = dict(error=repr(fail), key=key, arrival=requests[i][0], start=now, done=now,
                                  latency=now - requests[i][0], slot=None)
                    pending.popleft()
                    continue
                if slot is None:
                    self.lock.wait(timeout=0.05)
                    continue
                pending.popleft()
                th = threading.Thread(target=work, args=(i, slot))
                th.start()
                workers.append(th)
        for th in workers:
            th.join()
        return out, time.perf_counter() - t0


def poisson_arrivals(lam, count, keys, seed):
    """Seeded exponential gaps (rate lam per second) and uniformly drawn molecule keys."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / lam, size=count))
    pick = rng.integers(0, len(keys), size=count)
    return [(float(a), keys[j]) for a, j in zip(t, pick)]


def summarize(records, wall):
    lat = np.array([r["latency"] for r in records if "error" not in r])
    return dict(requests=len(records), errors=sum("error" in r for r in records), wall_s=round(wall, 3),
                molecules_per_s=round(len(lat) / wall, 2), p50_ms=round(1e3 * float(np.percentile(lat, 50)), 2),
                p99_ms=round(1e3 * float(np.percentile(lat, 99)), 2), mean_ms=round(1e3 * float(lat.mean()), 2))
