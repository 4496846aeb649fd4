"""Liquid crystal observables (reference: theories/liquid_crystal/observables.py)."""
from ...core.simulation import as_simulation, which_argument
from .kernels import CHARGE_KINDS


def compute_energy(sim, *args) -> float:
    sim = as_simulation(sim)
    return sim.compute_energy()


def compute_skyrmion_number(sim, *args) -> float:
    sim = as_simulation(sim)
    return sim.compute_charge(CHARGE_KINDS["skyrmion_number"])


def compute_electric_charge(sim, *args) -> float:
    """Bound flexoelectric charge -(div n div n + n.grad div n - (e3/e1)(...)) gv summed."""
    sim = as_simulation(sim)
    return sim.compute_charge(CHARGE_KINDS["electric_charge"])
