"""Seeded synthetic water clusters (H2O)_k used by the tests and the benchmark.

Monomer: O (0,0,0), H (0, +-0.757, 0.587) Angstrom.  Monomer m of (H2O)_k is rigidly rotated about its
oxygen by a uniform random rotation (Shoemake quaternion from three uniforms), optionally distorted
(bond +-db Angstrom, HOH angle +-da degrees), and placed at the m-th point of a 3.0 Angstrom cubic lattice
(points of a ceil(k^(1/3))^3 block in (x,y,z) lexicographic order) plus a uniform +-0.2 Angstrom jitter
per coordinate.  Random numbers come from splitmix64 with seed 1000 + k unless given.
"""
import math

import numpy as np

from .system import MolecularSystem

BOHR_TO_ANGSTROM = 0.529177210903  # CODATA 2018, as the .xyz reader of libzinq_ints.so
MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))


def _rotation(u1, u2, u3):
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    x, y, z, w = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2), b * math.sin(2 * math.pi * u3), b * math.cos(
        2 * math.pi * u3)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(k, seed=None, bond_jitter=0.0, angle_jitter=0.0, spacing=3.0, jitter=0.2):
    """Atomic numbers and Angstrom coordinates of (H2O)_k."""
    rng = SplitMix64(1000 + k if seed is None else seed)
    m = 1
    while m ** 3 < k:
        m += 1
    points = [(x, y, z) for x in range(m) for y in range(m) for z in range(m)][:k]
    r0 = math.hypot(0.757, 0.587)
    half0 = math.atan2(0.757, 0.587)
    Z, xyz = [], []
    for (px, py, pz) in points:
        R = _rotation(rng.uniform(), rng.uniform(), rng.uniform())
        r = r0 + bond_jitter * (2 * rng.uniform() - 1) if bond_jitter else r0
        half = half0 + math.radians(angle_jitter) * (2 * rng.uniform() - 1) / 2 if angle_jitter else half0
        mono = np.array([[0.0, 0.0, 0.0], [0.0, r * math.sin(half), r * math.cos(half)],
                         [0.0, -r * math.sin(half), r * math.cos(half)]])
        shift = np.array([px, py, pz]) * spacing + np.array([(2 * rng.uniform() - 1) * jitter for _ in range(3)])
        for atom, pos in zip((8, 1, 1), mono @ R.T + shift):
            Z.append(atom)
            xyz.append(pos)
    return np.array(Z, dtype=np.int32), np.array(xyz)


def to_xyz(Z, xyz, comment=""):
    sym = {1: "H", 8: "O"}
    lines = [str(len(Z)), comment] + ["%s %.10f %.10f %.10f" % (sym[z], *p) for z, p in zip(Z, xyz)]
    return "\n".join(lines) + "\n"


def water_system(k, basis="builtin:cc-pvdz", **kw):
    Z, xyz = water_cluster(k, **kw)
    return MolecularSystem.from_atoms(Z, xyz / BOHR_TO_ANGSTROM, basis)


BATCH_BASES = ("builtin:sto-3g", "builtin:cc-pvdz")


def batch_specs(count=512, seed=4242, kmax=3):
    """Seeded molecule list of the batched workload: k ~ U{1..kmax} waters, basis ~ U{STO-3G, cc-pVDZ},
    each monomer with +-0.05 Angstrom bond and +-5 degree angle distortions (own seed per molecule)."""
    rng = SplitMix64(seed)
    out = []
    for i in range(count):
        k = 1 + min(kmax - 1, int(rng.uniform() * kmax))
        basis = BATCH_BASES[min(1, int(rng.uniform() * 2))]
        out.append(dict(k=k, basis=basis, seed=int(rng.next() & 0xFFFFFFFF)))
    return out


def batch_system(spec):
    return water_system(spec["k"], basis=spec["basis"], seed=spec["seed"], bond_jitter=0.05, angle_jitter=5.0)
