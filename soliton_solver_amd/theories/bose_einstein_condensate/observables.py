"""Bose-Einstein condensate observables (reference: bose_einstein_condensate/observables.py:14-96)."""
from .kernels import CHARGE_KINDS


def compute_energy(sim) -> float:
    return sim.compute_energy()


def compute_norm(sim) -> float:
    """sum |psi|^2 gv."""
    return sim.compute_charge(CHARGE_KINDS["norm"])
