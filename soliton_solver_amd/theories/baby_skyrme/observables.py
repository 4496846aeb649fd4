"""Baby Skyrme observables (reference: theories/baby_skyrme/observables.py)."""
from ...core.simulation import as_simulation, which_argument
from .kernels import CHARGE_KINDS


def compute_energy(sim, *args) -> float:
    sim = as_simulation(sim)
    return sim.compute_energy()


def compute_skyrmion_number(sim, *args) -> float:
    sim = as_simulation(sim)
    return sim.compute_charge(CHARGE_KINDS["skyrmion_number"])
