"""Liquid crystal observables (reference: theories/liquid_crystal/observables.py)."""
from .kernels import CHARGE_KINDS


def compute_energy(sim) -> float:
    return sim.compute_energy()


def compute_skyrmion_number(sim) -> float:
    return sim.compute_charge(CHARGE_KINDS["skyrmion_number"])


def compute_electric_charge(sim) -> float:
    """Bound flexoelectric charge -(div n div n + n.grad div n - (e3/e1)(...)) gv summed."""
    return sim.compute_charge(CHARGE_KINDS["electric_charge"])
