"""Spin-triplet superconducting ferromagnet plugin (reference:
soliton_solver/theories/spin_triplet_superconducting_magnet/__init__.py)."""
from __future__ import annotations

from ..registry import TheorySpec
from . import initial_config, io, kernels, observables, params

THEORY_SPEC = TheorySpec(
    name="Spin triplet superconducting ferromagnet",
    import_path="soliton_solver_amd.theories.spin_triplet_superconducting_magnet",
    description="Ferromagnetic superconductor with equal-spin triplet pairing and mixed solitons, including "
                "skyrmions and multi-component vortices.",
    version="1.0",
)


def describe() -> None:
    print(f"{THEORY_SPEC.name} (version {THEORY_SPEC.version})")
    print(THEORY_SPEC.description)
    params.describe()
