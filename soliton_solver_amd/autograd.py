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

Parameters can be inputs of the same Function: ``discrete_energy(field, sim,
theta, names)`` takes a float64 tensor of named user parameters (e.g.
``["beta1", "beta2"]``), and its backward returns dE_h/dtheta as well.  That
gradient is exact: the ``ss_energy_param_grad`` kernel differentiates the
energy density with respect to every packed coupling constant p_f[j] with
forward-mode dual numbers (through the theory's own derivation of its
couplings from p_f), and the chain to the user parameters runs through the
host's ResolvedParams derivation (``packed_parameter_jacobian``).  At a
relaxed minimum phi* (dE_h/dphi = 0) it is dE*/dtheta (envelope theorem):
``energy_parameter_gradient``.
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


def packed_gradient(sim, field: torch.Tensor) -> torch.Tensor:
    """dE_h/dp_f (one entry per packed p_f value; the lattice-geometry entries
    0..5 are 0) at ``field``, from ``ss_energy_param_grad``."""
    f = _check_field(sim, field)
    n = int(sim.p_f_h.size)
    out = torch.empty(n, dtype=torch.float64, device=f.device)
    _lib.check(_lib.lib().ss_energy_param_grad(sim._ctx, C.c_void_p(f.data_ptr()), C.c_void_p(out.data_ptr()), n,
                                               C.c_void_p(_stream_ptr(f.device))), sim._ctx)
    return out


def packed_parameter_jacobian(sim, names, rel_step: float = 1e-6) -> np.ndarray:
    """J[j, k] = dp_f[j] / dtheta_k of the theory's packing
    (``pack_device_params(params.resolved())``) at the simulation's parameters,
    host-side: entries that copy a parameter are exactly 1 (the central
    difference of an identity, divided by the step actually taken: vp - vm is
    exact, the two values being within a factor 2), derived ones (e.g. the
    anisotropic theory's u1, u2 from alpha, beta:
    anisotropic_superconductor/params.py:190-243) are the central difference
    of that closed-form algebra at step rel_step |theta_k| (truncation ~
    step^2, rounding ~ 1e-16/step: ~1e-10 relative)."""
    base = sim.p
    th = sim.theory
    J = np.zeros((int(sim.p_f_h.size), len(names)))
    for k, name in enumerate(names):
        v = float(getattr(base, name))
        h = rel_step * max(abs(v), 1.0)
        vp, vm = v + h, v - h
        _, pf_p = th.params.pack_device_params(base.with_(**{name: vp}).resolved())
        _, pf_m = th.params.pack_device_params(base.with_(**{name: vm}).resolved())
        J[:, k] = (np.asarray(pf_p, dtype=np.float64) - np.asarray(pf_m, dtype=np.float64)) / (vp - vm)
    J[:6, :] = 0.0  # the lattice geometry does not depend on the couplings
    return J


class DiscreteEnergy(torch.autograd.Function):
    """``E = DiscreteEnergy.apply(field, sim, theta, names)``; ``E.backward()``
    gives dE_h/dfield and, when ``theta`` (a float64 tensor of the values of
    the user parameters ``names``, equal to the simulation's own) requires
    grad, dE_h/dtheta."""

    @staticmethod
    def forward(ctx, field, sim, theta=None, names=None):
        ctx.sim = sim
        ctx.names = list(names) if names is not None else []
        if theta is not None:
            cur = [float(getattr(sim.p, n)) for n in ctx.names]
            if theta.numel() != len(cur) or any(abs(a - b) > 1e-15 * max(1.0, abs(b))
                                                for a, b in zip(theta.detach().cpu().tolist(), cur)):
                raise ValueError("theta must hold the simulation's current values of the named parameters "
                                 "(set them with sim.set_params)")
        ctx.save_for_backward(field)
        return energy_of(sim, field)

    @staticmethod
    def backward(ctx, grad_out):
        (field,) = ctx.saved_tensors
        g_field = None
        if ctx.needs_input_grad[0]:
            g_field = grad_out * exact_gradient(ctx.sim, field.detach())
        g_theta = None
        if len(ctx.needs_input_grad) > 2 and ctx.needs_input_grad[2]:
            dpf = packed_gradient(ctx.sim, field.detach()).cpu().numpy()
            J = packed_parameter_jacobian(ctx.sim, ctx.names)
            g_theta = grad_out * torch.tensor(J.T @ dpf, dtype=torch.float64, device=grad_out.device)
        return g_field, None, g_theta, None


def discrete_energy(field: torch.Tensor, sim, theta: torch.Tensor | None = None, names=None) -> torch.Tensor:
    return DiscreteEnergy.apply(field, sim, theta, names)


def energy_parameter_gradient(sim, field: torch.Tensor, names):
    """dE_h/dtheta at fixed ``field`` for the named user parameters, by the
    backward of ``DiscreteEnergy``.  Dependent packed values (e.g. the
    anisotropic theory's u1, u2, theta_diff, beta3tilde) are chained through
    the theory's own ``ResolvedParams``, so the derivative includes their
    dependence on theta.  At a relaxed minimum phi*, dE_h/dphi = 0, so this is
    dE*/dtheta (envelope theorem).  Returns {name: value}."""
    theta = torch.tensor([float(getattr(sim.p, n)) for n in names], dtype=torch.float64, device=field.device,
                         requires_grad=True)
    E = discrete_energy(field.detach(), sim, theta, names)
    (g,) = torch.autograd.grad(E, theta)
    return {n: float(v) for n, v in zip(names, g.cpu().tolist())}


__all__ = ["DiscreteEnergy", "discrete_energy", "energy_of", "exact_gradient", "reference_variation",
           "packed_gradient", "packed_parameter_jacobian", "energy_parameter_gradient"]
