"""Anyon superconductor (Maxwell-Chern-Simons-Higgs) plugin (reference:
soliton_solver/theories/maxwell_chern_simons_higgs/__init__.py)."""
from __future__ import annotations

from ..registry import TheorySpec
from . import initial_config, io, kernels, observables, params

THEORY_SPEC = TheorySpec(
    name="Anyon superconductor",
    aliases=("Maxwell Chern-Simons Higgs", "Chern-Simons-Landau-Ginzburg", "Abelian Chern-Simons Higgs"),
    import_path="soliton_solver_amd.theories.maxwell_chern_simons_higgs",
    description="Vortex anyons in Maxwell-Chern-Simons-Higgs or Chern-Simons-Landau-Ginzburg models.",
    version="1.0",
)


def describe() -> None:
    print(f"{THEORY_SPEC.name} (version {THEORY_SPEC.version})")
    print(THEORY_SPEC.description)
    params.describe()
