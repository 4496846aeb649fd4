"""soliton_solver_amd: the MI355X (gfx950) relaxation core of soliton_solver.

Public API mirrors the reference: ``load_theory(name)``,
``Simulation(params, theory)``, ``sim.initialize(cfg)``, ``sim.step(E)``,
``sim.observables()``.  The arithmetic runs in hand-written HIP kernels in
``libsoliton.so`` (built by ``python -m soliton_solver_amd.build``).
"""
from __future__ import annotations

__version__ = "0.1.0"

from .core.simulation import Simulation  # noqa: E402
from .theories import list_theories, load_theory  # noqa: E402

__all__ = ["Simulation", "load_theory", "list_theories", "__version__"]
