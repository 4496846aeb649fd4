"""Ginzburg-Landau superconductor / Abelian Higgs plugin (reference:
soliton_solver/theories/ginzburg_landau_superconductor/__init__.py)."""
from __future__ import annotations

from ..registry import TheorySpec
from . import initial_config, io, kernels, observables, params

THEORY_SPEC = TheorySpec(
    name="Ginzburg-Landau superconductor",
    aliases=("Abelian Higgs", "Ginzburg-Landau", "Abelian Higgs model", "Ginzburg-Landau model"),
    import_path="soliton_solver_amd.theories.ginzburg_landau_superconductor",
    description="Superconducting vortices in the Ginzburg-Landau or Abelian Higgs model.",
    version="1.0",
)


def describe() -> None:
    print(f"{THEORY_SPEC.name} (version {THEORY_SPEC.version})")
    print(THEORY_SPEC.description)
    params.describe()
