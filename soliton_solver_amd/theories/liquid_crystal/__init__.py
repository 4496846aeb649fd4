"""Chiral liquid crystal plugin (reference: soliton_solver/theories/liquid_crystal/__init__.py)."""
from __future__ import annotations

from ..registry import TheorySpec
from . import initial_config, io, kernels, observables, params

THEORY_SPEC = TheorySpec(
    name="Liquid crystal",
    aliases=("Chiral liquid crystal", "Liquid crystal skyrmion", "Nematic liquid crystal"),
    import_path="soliton_solver_amd.theories.liquid_crystal",
    description="Skyrmions in liquid crystal systems with twist or splay-bend preference and flexoelectric "
                "depolarization.",
    version="1.1",
)


def describe() -> None:
    print(f"{THEORY_SPEC.name} (version {THEORY_SPEC.version})")
    print(THEORY_SPEC.description)
    params.describe()
