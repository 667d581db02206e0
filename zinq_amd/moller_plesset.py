"""MP2 on the GPU transform: mirror of src/moller_plesset.zig run / runFromSystem (:115-262).

For order 2 the reference's determinant-space RS perturbation theory (mp, :340-481) reduces exactly to
the closed-shell sum over (ia|jb) (singles couple with zero because H_MS is built from the orbital
energies, :377-388), which zq_mp2 evaluates from the OVOV transform in FP64.  Higher orders need the
determinant machinery, which is not part of the GPU hot path.
"""
from dataclasses import dataclass, field
from typing import List

from . import hartree_fock, hip
from .system import MolecularSystem


@dataclass
class Options:
    hartree_fock: hartree_fock.Options
    order: int = 2
    precision: str = "fp64"  # quarter transforms: fp64 | fp32 | bf16x3


@dataclass
class Result:
    hartree_fock: hartree_fock.Result
    energy: List[float]
    correlation: float = 0.0
    context: object = field(repr=False, default=None)


def run_from_system(opt, sys, ctx=None):
    if opt.order != 2:
        raise hip.InvalidInput("only order 2 runs on the GPU transform")
    if opt.hartree_fock.generalized:
        raise hip.InvalidInput("the GPU MP2 path takes a restricted (RHF) reference")
    hf = hartree_fock.run_from_system(opt.hartree_fock, sys, ctx)
    e2 = hf.context.mp2(hf.C, hf.e, hf.nocc, opt.precision)
    return Result(hartree_fock=hf, energy=[hf.energy[0] + e2], correlation=e2, context=hf.context)


def run(opt, ctx=None):
    h = opt.hartree_fock
    sys = MolecularSystem.from_xyz(h.system, h.basis, h.charge, h.multiplicity)
    return run_from_system(opt, sys, ctx)
