"""Chiral magnet theory plugin (reference: soliton_solver/theories/chiral_magnet/__init__.py:22-28)."""
from __future__ import annotations

from ..registry import TheorySpec
from . import initial_config, io, kernels, observables, params

THEORY_SPEC = TheorySpec(
    name="Chiral magnet",
    aliases=("Chiral ferromagnet", "Chiral magnetic skyrmion", "Magnetic skyrmion"),
    import_path="soliton_solver_amd.theories.chiral_magnet",
    description="Skyrmions in chiral ferromagnetic systems, with various DMI choices and optional demagnetization.",
    version="1.1",
)


def describe() -> None:
    print(f"{THEORY_SPEC.name} (version {THEORY_SPEC.version})")
    print(THEORY_SPEC.description)
    params.describe()
