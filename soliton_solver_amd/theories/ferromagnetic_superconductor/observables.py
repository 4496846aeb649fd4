"""Ferromagnetic superconductor observables (reference:
theories/ferromagnetic_superconductor/observables.py:13-130)."""
from ...core.simulation import as_simulation, which_argument
from .kernels import CHARGE_KINDS


def compute_energy(sim, *args) -> float:
    sim = as_simulation(sim)
    return sim.compute_energy()


def compute_skyrmion_number(sim, *args) -> float:
    """m.(d_x m x d_y m) gv / (4 pi M0^3) summed over the lattice."""
    sim = as_simulation(sim)
    return sim.compute_charge(CHARGE_KINDS["skyrmion_number"])


def compute_vortex_number(sim, *args, which: int = 2) -> float:
    """Flux component over 2 pi q: which 0 -> d_y A3, 1 -> -d_x A3, otherwise d_x A2 - d_y A1."""
    which = which_argument(args, which)
    sim = as_simulation(sim)
    key = {0: "vortex_number_0", 1: "vortex_number_1"}.get(int(which), "vortex_number")
    return sim.compute_charge(CHARGE_KINDS[key])
