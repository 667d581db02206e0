"""Hartree-Fock with the GPU Fock build: mirror of src/hartree_fock.zig (non-DFT energy path).

``Options`` keeps the reference's field names and defaults (src/hartree_fock.zig:52-105);
``run``/``run_from_system`` follow :315-509 (integrals, V_NN, closed-shell check, core guess, scf).
The SCF loop itself is the C++ driver zq_scf_run (include/zinq_scf.h) over zq_fock_jk.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import hip
from .system import MolecularSystem, spin_block_1e, spin_block_2e


class OnlyClosedShellSupported(ValueError):
    """error.OnlyClosedShellSupported, src/hartree_fock.zig:378."""


@dataclass
class Options:
    system: str
    basis: str
    diis: Optional[int] = 8
    generalized: bool = False
    charge: int = 0
    multiplicity: int = 1
    iterations: int = 100
    threshold: float = 1e-8


@dataclass
class Result:
    C: np.ndarray
    P: np.ndarray
    F: np.ndarray
    e: np.ndarray
    energy: List[float]
    iterations: int
    nocc: int
    H: np.ndarray = field(repr=False, default=None)
    S: np.ndarray = field(repr=False, default=None)
    VN: float = 0.0
    context: object = field(repr=False, default=None)


def integrals(sys, generalized):
    """H, S, dense g, V_NN of src/molecular_integrals.zig:152-204 (spin-blocked for GHF)."""
    S, H, g = sys.overlap(), sys.hcore(), sys.coulomb()
    if generalized:
        S, H, g = spin_block_1e(S), spin_block_1e(H), spin_block_2e(g)
    return H, S, g, sys.nrep()


def run_from_system(opt, sys, ctx=None):
    if opt.generalized:
        nocc = sys.nel
    else:
        if sys.nel % 2 != 0:
            raise OnlyClosedShellSupported("ONLY CLOSED SHELL SUPPORTED")
        nocc = sys.nel // 2
    ctx = ctx or hip.Context(0)
    if opt.generalized:
        H, S, g, VN = integrals(sys, True)
        ctx.eri_load(g)
    else:
        S, H, VN = sys.overlap(), sys.hcore(), sys.nrep()
        ctx.eri_load_packed(sys.coulomb_packed(), sys.nbf)
    r = ctx.scf(H, S, VN, nocc, diis=opt.diis, iterations=opt.iterations, threshold=opt.threshold,
                generalized=opt.generalized)
    return Result(C=r["C"], P=r["P"], F=r["F"], e=r["e"], energy=[r["E"]], iterations=r["iterations"], nocc=nocc, H=H,
                  S=S, VN=VN, context=ctx)


def run(opt, ctx=None):
    """src/hartree_fock.zig run: build the system from the .xyz file and basis, then run_from_system."""
    sys = MolecularSystem.from_xyz(opt.system, opt.basis, opt.charge, opt.multiplicity)
    return run_from_system(opt, sys, ctx)
