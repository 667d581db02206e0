import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs through the C ABI of libzinq_hip.so")


@pytest.fixture(scope="session")
def water_xyz():
    return os.path.join(GOLDEN, "water.xyz")


_cache = {}


def system_ints(kind):
    """(system, H, S, g_dense, VN) for 'sto3g', 'ccpvdz', or 'w2' / 'w3' ((H2O)_k, cc-pVDZ)."""
    if kind not in _cache:
        from zinq_amd.system import MolecularSystem
        from zinq_amd.molecules import water_system
        if kind == "sto3g":
            s = MolecularSystem.from_xyz(os.path.join(GOLDEN, "water.xyz"), "builtin:sto-3g")
        elif kind == "ccpvdz":
            s = MolecularSystem.from_xyz(os.path.join(GOLDEN, "water.xyz"), "builtin:cc-pvdz")
        else:
            s = water_system(int(kind[1:]))
        _cache[kind] = (s, s.hcore(), s.overlap(), s.coulomb(), s.nrep())
    return _cache[kind]
