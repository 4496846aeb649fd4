"""Anisotropic superconductor observables (reference: anisotropic_superconductor/observables.py)."""
from ...core.simulation import as_simulation, which_argument
from .kernels import CHARGE_KINDS


def compute_energy(sim, *args) -> float:
    sim = as_simulation(sim)
    return sim.compute_energy()


def compute_vortex_number(sim, *args, which: int = 2) -> float:
    which = which_argument(args, which)
    sim = as_simulation(sim)
    return sim.compute_charge(CHARGE_KINDS["vortex_number"])
