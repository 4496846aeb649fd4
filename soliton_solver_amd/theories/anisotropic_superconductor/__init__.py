"""Anisotropic s+id superconductor plugin (reference:
soliton_solver/theories/anisotropic_superconductor/__init__.py)."""
from __future__ import annotations

from ..registry import TheorySpec
from . import initial_config, io, kernels, observables, params

THEORY_SPEC = TheorySpec(
    name="Anisotropic superconductor",
    aliases=("s+id superconductor", "Multicomponent superconductor", "Multi-component superconductor",
             "Multicomponent GL", "Multi-component GL", "Multicomponent Ginzburg-Landau",
             "Multi-component Ginzburg-Landau"),
    import_path="soliton_solver_amd.theories.anisotropic_superconductor",
    description="Multi-component anisotropic s+id superconductor, simulating multi-component vortices.",
    version="1.0",
)


def describe() -> None:
    print(f"{THEORY_SPEC.name} (version {THEORY_SPEC.version})")
    print(THEORY_SPEC.description)
    params.describe()
