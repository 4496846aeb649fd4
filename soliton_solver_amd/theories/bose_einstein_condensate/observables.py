"""Bose-Einstein condensate observables (reference: bose_einstein_condensate/observables.py:14-96)."""
from ...core.simulation import as_simulation, which_argument
from .kernels import CHARGE_KINDS


def compute_energy(sim, *args) -> float:
    sim = as_simulation(sim)
    return sim.compute_energy()


def compute_norm(sim, *args) -> float:
    """sum |psi|^2 gv."""
    sim = as_simulation(sim)
    return sim.compute_charge(CHARGE_KINDS["norm"])
