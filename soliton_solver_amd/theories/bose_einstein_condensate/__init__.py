"""Bose-Einstein condensate plugin (reference: soliton_solver/theories/bose_einstein_condensate/__init__.py)."""
from __future__ import annotations

from ..registry import TheorySpec
from . import initial_config, io, kernels, observables, params

THEORY_SPEC = TheorySpec(
    name="Bose-Einstein condensate",
    aliases=("BEC", "Gross-Pitaevskii", "Gross-Pitaevskii equation", "GP", "GPE"),
    import_path="soliton_solver_amd.theories.bose_einstein_condensate",
    description="Vortices in a rotating Bose-Einstein condensate (BEC).",
    version="1.0",
)


def describe() -> None:
    print(f"{THEORY_SPEC.name} (version {THEORY_SPEC.version})")
    print(THEORY_SPEC.description)
    params.describe()
