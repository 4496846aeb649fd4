"""Chiral magnet observables (reference: theories/chiral_magnet/observables.py:15-141).

Each returns a Python float: the FP64 sum over the lattice of the density
times the cell area, reduced on the GPU."""
from ...core.simulation import as_simulation, which_argument
from .kernels import CHARGE_KINDS


def compute_energy(sim, *args) -> float:
    sim = as_simulation(sim)
    return sim.compute_energy()


def compute_skyrmion_number(sim, *args) -> float:
    sim = as_simulation(sim)
    return sim.compute_charge(CHARGE_KINDS["skyrmion_number"])


def compute_magnetic_charge(sim, *args) -> float:
    sim = as_simulation(sim)
    return sim.compute_charge(CHARGE_KINDS["magnetic_charge"])
