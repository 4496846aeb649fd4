"""The discrete energy as a ``torch.autograd.Function``.

Forward: E_h(phi) = sum_sites gv e(phi, D1 phi), the theory's energy summed on
the GPU in FP64 (the same kernel ``Simulation.compute_energy`` uses).
Backward: dE_h/dphi from the exact-adjoint variation kernel (forward-mode
dual numbers through the theory's energy density, then the transposed
first-derivative stencil), so the gradient is the exact gradient of the
discrete energy, not the reference's continuum Euler-Lagrange expression
(which differs from it by O(h^4) and by the reference's masks and factors).

PyTorch is used only at this boundary: tensors are FP64 CUDA tensors in the
reference layout ``(nf, xlen, ylen)``; the kernels run on torch's current
stream through ``ss_energy_dev`` / ``ss_variation_dev``.

Parameter sensitivities at a relaxed minimum phi* (dE*/dtheta = dE_h/dtheta
at phi*, by stationarity) are provided by ``energy_parameter_gradient``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib


def _stream_ptr(device) -> int:
    return int(torch.cuda.current_stream(device).cuda_stream)


def _check_field(sim, field: torch.Tensor) -> torch.Tensor:
    if not field.is_cuda or field.dtype != torch.float64:
        raise TypeError("field must be a float64 CUDA tensor")
    if tuple(field.shape) != tuple(sim.field_shape):
        raise ValueError(f"field shape {tuple(field.shape)} != {tuple(sim.field_shape)}")
    if sim.precision != "fp64":
        raise ValueError("the autograd boundary needs an FP64 Simulation")
    return field.contiguous()


def energy_of(sim, field: torch.Tensor) -> torch.Tensor:
    """E_h(field) as a 0-d float64 tensor (no autograd)."""
    f = _check_field(sim, field)
    out = torch.empty(1, dtype=torch.float64, device=f.device)
    _lib.check(_lib.lib().ss_energy_dev(sim._ctx, C.c_void_p(f.data_ptr()), C.c_void_p(out.data_ptr()),
                                        C.c_void_p(_stream_ptr(f.device))), sim._ctx)
    return out[0]


def exact_gradient(sim, field: torch.Tensor) -> torch.Tensor:
    """dE_h/dphi of ``field`` from the exact-adjoint kernel."""
    f = _check_field(sim, field)
    g = torch.empty_like(f)
    _lib.check(_lib.lib().ss_variation_dev(sim._ctx, C.c_void_p(f.data_ptr()), C.c_void_p(g.data_ptr()),
                                           _lib.VAR_EXACT_ADJOINT, C.c_void_p(_stream_ptr(f.device))), sim._ctx)
    return g


def reference_variation(sim, field: torch.Tensor) -> torch.Tensor:
    """The reference's energy gradient EG (the flow's variation) of ``field``."""
    f = _check_field(sim, field)
    g = torch.empty_like(f)
    _lib.check(_lib.lib().ss_variation_dev(sim._ctx, C.c_void_p(f.data_ptr()), C.c_void_p(g.data_ptr()),
                                           _lib.VAR_REFERENCE, C.c_void_p(_stream_ptr(f.device))), sim._ctx)
    return g


class DiscreteEnergy(torch.autograd.Function):
    """``E = DiscreteEnergy.apply(field, sim)``; ``E.backward()`` gives dE_h/dfield."""

    @staticmethod
    def forward(ctx, field, sim):
        ctx.sim = sim
        ctx.save_for_backward(field)
        return energy_of(sim, field)

    @staticmethod
    def backward(ctx, grad_out):
        (field,) = ctx.saved_tensors
        g = exact_gradient(ctx.sim, field.detach())
        return grad_out * g, None


def discrete_energy(field: torch.Tensor, sim) -> torch.Tensor:
    return DiscreteEnergy.apply(field, sim)


def energy_parameter_gradient(sim, field: torch.Tensor, names, rel_step: float = 1e-3):
    """dE_h/dtheta at fixed ``field`` for named theory parameters.

    Dependent packed values (e.g. the anisotropic theory's u1, u2, theta_diff,
    beta3tilde) are re-derived by the theory's own ``ResolvedParams``, so the
    derivative includes their dependence on theta.  It is the 5-point central
    difference in theta of the FP64 GPU energy at fixed field, step
    ``rel_step*|theta|`` (truncation O(step^4), about 1e-12 relative for the
    smooth theta-dependence of these energies).  At a relaxed minimum phi*,
    dE_h/dphi = 0, so this is dE*/dtheta (envelope theorem).
    """
    base = sim.p
    out = {}
    for name in names:
        v = float(getattr(base, name))
        h = rel_step * max(abs(v), 1e-3)
        vals = []
        for k in (-2, -1, 1, 2):
            sim.set_params(base.with_(**{name: v + k * h}))
            vals.append(float(energy_of(sim, field)))
        sim.set_params(base)
        out[name] = (vals[0] - 8.0 * vals[1] + 8.0 * vals[2] - vals[3]) / (12.0 * h)
    return out


__all__ = ["DiscreteEnergy", "discrete_energy", "energy_of", "exact_gradient", "reference_variation",
           "energy_parameter_gradient"]
