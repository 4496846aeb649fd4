"""Anisotropic superconductor observables (reference: anisotropic_superconductor/observables.py)."""
from .kernels import CHARGE_KINDS


def compute_energy(sim) -> float:
    return sim.compute_energy()


def compute_vortex_number(sim, which: int = 2) -> float:
    return sim.compute_charge(CHARGE_KINDS["vortex_number"])
