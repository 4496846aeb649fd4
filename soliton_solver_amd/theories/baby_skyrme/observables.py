"""Baby Skyrme observables (reference: theories/baby_skyrme/observables.py)."""
from .kernels import CHARGE_KINDS


def compute_energy(sim) -> float:
    return sim.compute_energy()


def compute_skyrmion_number(sim) -> float:
    return sim.compute_charge(CHARGE_KINDS["skyrmion_number"])
