"""Baby Skyrme model plugin (reference: soliton_solver/theories/baby_skyrme/__init__.py)."""
from __future__ import annotations

from ..registry import TheorySpec
from . import initial_config, io, kernels, observables, params

THEORY_SPEC = TheorySpec(
    name="Baby Skyrme model",
    aliases=("Baby Skyrme", "Baby skyrmion", "Planar skyrmion"),
    import_path="soliton_solver_amd.theories.baby_skyrme",
    description="Skyrmions in a 2D planar Skyrme model, with numerous choices of potentials.",
    version="1.1",
)


def describe() -> None:
    print(f"{THEORY_SPEC.name} (version {THEORY_SPEC.version})")
    print(THEORY_SPEC.description)
    params.describe()
