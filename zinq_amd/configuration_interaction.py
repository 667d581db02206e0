"""Configuration interaction on the GPU: mirror of src/configuration_interaction.zig run (energies).

``Options`` keeps the reference's field names and defaults for the energy path (:51-81: hartree_fock,
excitations = [1, 2]); gradients, optimisation and Hessians are outside the hot path.  The states come
from zq_ci: the FULL AO->MO transform on the GPU, the Slater-Condon H_CI over generateDets
(:126-183, :323-432) built on the GPU, and its eigenvalues + V_NN (computeCiStates, :597-683).
"""
from dataclasses import dataclass, field
from typing import List

from . import hartree_fock, hip
from .system import MolecularSystem


@dataclass
class Options:
    hartree_fock: hartree_fock.Options
    excitations: List[int] = field(default_factory=lambda: [1, 2])
    nstates: int = 5


@dataclass
class Result:
    hartree_fock: hartree_fock.Result
    energy: List[float]
    ndet: int = 0
    context: object = field(repr=False, default=None)


def run_from_system(opt, sys, ctx=None):
    hf = hartree_fock.run_from_system(opt.hartree_fock, sys, ctx)
    E, nd = hf.context.ci(hf.H, hf.C, hf.nocc, opt.excitations, hf.VN, opt.nstates, opt.hartree_fock.generalized)
    return Result(hartree_fock=hf, energy=[float(x) for x in E], ndet=nd, context=hf.context)


def run(opt, ctx=None):
    h = opt.hartree_fock
    sys = MolecularSystem.from_xyz(h.system, h.basis, h.charge, h.multiplicity)
    return run_from_system(opt, sys, ctx)
