"""Moller-Plesset on the GPU transforms: mirror of src/moller_plesset.zig run / runFromSystem (:115-262).

For order 2 with a restricted reference the determinant-space RS perturbation theory (mp, :340-481)
reduces exactly to the closed-shell sum over (ia|jb) (singles couple with zero because H_MS is built from
the orbital energies, :377-388), which zq_mp2 evaluates from the OVOV transform in FP64.  Higher orders
and generalized (GHF) references run the determinant recursion itself on the GPU (zq_mp) over the FULL
transform.  energy[0] is the total energy E_HF + sum of the order >= 2 corrections, as the reference's.
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
    if opt.order < 1:
        raise hip.InvalidInput("order must be at least 1")
    hf = hartree_fock.run_from_system(opt.hartree_fock, sys, ctx)
    if opt.order == 2 and not opt.hartree_fock.generalized:
        corr = hf.context.mp2(hf.C, hf.e, hf.nocc, opt.precision)
    else:
        if opt.precision != "fp64":
            raise hip.InvalidInput("the determinant path runs its FULL transform in FP64")
        E, _ = hf.context.mp(hf.C, hf.e, hf.nocc, opt.order, opt.hartree_fock.generalized)
        corr = float(E[2:].sum())
    return Result(hartree_fock=hf, energy=[hf.energy[0] + corr], correlation=corr, context=hf.context)


def run(opt, ctx=None):
    h = opt.hartree_fock
    sys = MolecularSystem.from_xyz(h.system, h.basis, h.charge, h.multiplicity)
    return run_from_system(opt, sys, ctx)
