"""Host logic of the serving loop (zinq_amd/serve.py) on the CPU: admission by footprint, LRU eviction of
idle ERI sets, continuous batching over the slots.  The GPU calls are replaced by a stand-in module that
records what the server does with them; the GPU tests run the same loop on the device."""
import threading
import time

import pytest

from zinq_amd import serve as sv


class _Mol:
    def __init__(self, key, n):
        self.key, self.n, self.o = key, n, n // 8
        self.H = self.S = self.packed = None
        self.VN = 0.0


def _fake_hip(budget_box, log):
    lock = threading.Lock()
    live = set()
    running = [0]

    class EriSet:
        def __init__(self, packed, n, device, with_fp32):
            self.n, self.closed = n, False
            with lock:
                used = sum(sv.eri_set_bytes(s.n, with_fp32) for s in live) + sv.eri_set_bytes(n, with_fp32)
                if used > budget_box[0]:
                    log.append(f"over budget: {used} > {budget_box[0]}")
                live.add(self)
                log.append(f"create {n}")

        def close(self):
            with lock:
                live.discard(self)
                self.closed = True
                log.append(f"close {self.n}")

    class Context:
        def __init__(self, device):
            self.set = None

        def attach(self, s):
            if s.closed:
                raise RuntimeError("attached a closed set")
            self.set = s

        def scf(self, H, S, VN, nocc, **kw):
            with lock:
                running[0] += 1
                log.append(("running", running[0]))
            time.sleep(0.003)
            closed = self.set.closed
            with lock:
                running[0] -= 1
            if closed:
                raise RuntimeError("set freed while in use")
            return dict(E=-float(self.set.n), C=None, e=None, iterations=3)

        def mp2(self, C, e, nocc, precision):
            if self.set.closed:
                raise RuntimeError("set freed while in use")
            return -1e-3 * self.set.n

        def mp2_resident(self, nocc, precision):
            return self.mp2(None, None, nocc, precision)

        def close(self):
            pass

    class Hip:
        pass

    Hip.EriSet, Hip.Context, Hip.GpuOutOfMemory = EriSet, Context, MemoryError
    from zinq_amd import hip as real_hip  # host arithmetic of the library, no device needed
    Hip.jk_record_count = staticmethod(real_hip.jk_record_count)
    return Hip


@pytest.mark.parametrize("slots", [1, 3])
def test_admission_eviction_and_slots(monkeypatch, slots):
    mols = [_Mol(f"m{n}", n) for n in (48, 56, 64, 72)]  # sets of 11, 20, 33 and 51 MB
    reserve = max(sv.scratch_bytes(m.n, m.o, 4) for m in mols)
    big = max(sv.eri_set_bytes(m.n, True) for m in mols)
    budget_box, log = [0], []
    monkeypatch.setattr(sv, "hip", _fake_hip(budget_box, log))
    hbm = slots * reserve + int(1.3 * big)  # the largest set and a small one at most
    srv = sv.Server(mols, slots=slots, hbm_budget=hbm)
    budget_box[0] = srv.budget
    keys = [m.key for m in mols]
    rec, wall = srv.serve(sv.poisson_arrivals(400.0, 40, keys, seed=9))
    srv.close()
    assert len(rec) == 40 and wall > 0
    for r in rec:
        assert "error" not in r, r
        n = int(r["key"][1:])
        assert r["E_rhf"] == -float(n) and r["E_corr"] == pytest.approx(-1e-3 * n)  # its own set
        assert r["latency"] >= r["done"] - r["start"] >= 0
    assert not [x for x in log if isinstance(x, str) and x.startswith("over budget")]
    assert max(x[1] for x in log if isinstance(x, tuple)) <= slots
    creates = sum(1 for x in log if isinstance(x, str) and x.startswith("create"))
    assert creates > len(mols)  # the budget forced evictions and re-creations
    assert srv.reserved <= srv.budget


def test_molecule_above_the_budget_fails_instead_of_waiting(monkeypatch):
    """A set larger than the whole budget can never be admitted: its requests fail at once with an
    out-of-memory record, and the others are served."""
    mols = [_Mol("m24", 24), _Mol("m72", 72)]
    reserve = max(sv.scratch_bytes(m.n, m.o, 4) for m in mols)
    budget_box, log = [0], []
    monkeypatch.setattr(sv, "hip", _fake_hip(budget_box, log))
    srv = sv.Server(mols, slots=2, hbm_budget=2 * reserve + 2 * sv.eri_set_bytes(24, True))
    budget_box[0] = srv.budget
    rec, _ = srv.serve([(0.0, "m72"), (0.001, "m24"), (0.002, "m72")])
    srv.close()
    assert [("error" in r) for r in rec] == [True, False, True]
    assert "budget" in rec[0]["error"] and rec[1]["E_rhf"] == -24.0


def test_unknown_molecule_and_failed_set_are_recorded(monkeypatch):
    """A request naming no known molecule, and one whose ERI set cannot be created, each get an error
    record; the loop keeps serving the rest."""
    mols = [_Mol("m24", 24), _Mol("m32", 32)]
    hipmod = _fake_hip([10 ** 15], [])
    make = hipmod.EriSet

    class FailingSet(make):
        def __init__(self, packed, n, device, with_fp32):
            if n == 32:
                raise MemoryError("set allocation failed")
            super().__init__(packed, n, device, with_fp32)

    hipmod.EriSet = FailingSet
    monkeypatch.setattr(sv, "hip", hipmod)
    srv = sv.Server(mols, slots=2, hbm_budget=10 ** 12)
    rec, _ = srv.serve([(0.0, "nope"), (0.001, "m32"), (0.002, "m24")])
    srv.close()
    assert "unknown molecule" in rec[0]["error"] and "allocation failed" in rec[1]["error"]
    assert "error" not in rec[2] and rec[2]["E_rhf"] == -24.0


def test_budget_below_the_slot_reserve_is_refused(monkeypatch):
    mols = [_Mol("m72", 72)]
    monkeypatch.setattr(sv, "hip", _fake_hip([0], []))
    with pytest.raises(MemoryError):
        sv.Server(mols, slots=4, hbm_budget=sv.scratch_bytes(72, 9, 4))
