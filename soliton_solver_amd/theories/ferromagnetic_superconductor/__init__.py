"""Ferromagnetic superconductor plugin (reference:
soliton_solver/theories/ferromagnetic_superconductor/__init__.py)."""
from __future__ import annotations

from ..registry import TheorySpec
from . import initial_config, io, kernels, observables, params

THEORY_SPEC = TheorySpec(
    name="Ferromagnetic superconductor",
    aliases=("Superconducting ferromagnet", "Superconducting magnet", "Magnetic skyrmions superconducting vortices",
             "skyrmion vortex pairs"),
    import_path="soliton_solver_amd.theories.ferromagnetic_superconductor",
    description="Hybrid magnetic skyrmion and superconducting vortices in ferromagnetic superconductors.",
    version="1.1",
)


def describe() -> None:
    print(f"{THEORY_SPEC.name} (version {THEORY_SPEC.version})")
    print(THEORY_SPEC.description)
    params.describe()
