"""Handles of the native (HIP) kernels a theory plugin exposes.

The reference builds its gradient and RK4 kernels with factories
(``make_do_gradient_step_kernel`` core/integrator.py:15-48,
``make_do_rk4_kernel`` :151-189) that close over a theory's device functions.
Here the kernels are compiled into libsoliton.so per theory; a theory's
``kernels`` module names them with ``NativeKernel`` handles so that the
reference's plugin checks (``do_gradient_step_kernel`` required,
``do_rk4_kernel`` optional, core/simulation.py:74-80) keep their meaning.
"""
from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class NativeKernel:
    """Name of one kernel family of a theory inside libsoliton.so."""

    theory_id: int
    role: str           # "gradient", "rk4_constrained", "energy", "charge:<name>"

    def __call__(self, *args, **kwargs):
        raise TypeError("native kernels are launched by Simulation through libsoliton.so, not called directly")


def make_do_gradient_step_kernel(theory_id: int) -> NativeKernel:
    return NativeKernel(theory_id, "gradient")


def make_do_rk4_kernel(theory_id: int) -> NativeKernel:
    return NativeKernel(theory_id, "rk4_constrained")
