"""Maxwell-Chern-Simons-Higgs observables (reference: maxwell_chern_simons_higgs/observables.py:14-190)."""
from .kernels import CHARGE_KINDS


def compute_energy(sim) -> float:
    return sim.compute_energy()


def compute_vortex_number(sim, which: int = 2) -> float:
    return sim.compute_charge(CHARGE_KINDS["vortex_number"])


def compute_electric_charge(sim) -> float:
    return sim.compute_charge(CHARGE_KINDS["electric_charge"])


def compute_noether_charge(sim) -> float:
    return sim.compute_charge(CHARGE_KINDS["noether_charge"])
