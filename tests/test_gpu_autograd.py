"""Exact-adjoint variation and the torch.autograd boundary (GPU).

* dE_h/dphi from the dual-number adjoint kernel vs torch autograd of the
  oracle's pure-tensor energy on the same field (FP64, 1e-11 of max).
* torch.autograd.gradcheck of DiscreteEnergy against central differences
  (eps = 1e-6) on a 32x32 patch of a 512x512 anisotropic s+id field, the rest
  frozen.
* The exact-adjoint flow relaxes to a stationary point of E_h.
* dE*/dbeta_i at a relaxed minimum vs re-relaxation at beta_i +- delta.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import soliton_solver_amd as ssa
from oracle import core as ocore
from soliton_solver_amd import autograd as sag
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _torch_grad(o, F, p_i, p_f):
    Ft = torch.tensor(F, dtype=torch.float64, requires_grad=True)
    E = o.energy_density(Ft, ocore.first_derivatives(Ft, p_i, p_f), p_i, p_f).sum()
    E.backward()
    return float(E.detach()), Ft.grad.numpy()


CASES = [
    ("Chiral magnet", lambda th: H.chiral_params(th, 40, 52), lambda th, pi, pf: H.chiral_field(th, pi, pf, seed=2, rough=True)),
    ("Ginzburg-Landau superconductor", lambda th: H.gl_params(th, 44, 38), lambda th, pi, pf: H.gl_field(th, pi, pf, seed=1)),
    ("Baby Skyrme model", lambda th: H.bs_params(th, 50, 46, potential="broken", N=2, kappa=1.2),
     lambda th, pi, pf: H.bs_field(th, pi, pf, seed=3, k=2, amp=1e-2)),
    ("Anisotropic superconductor", lambda th: H.as_params(th, 48, 40, seed=4),
     lambda th, pi, pf: H.as_field(th, pi, pf, seed=4, amp=5e-2)),
    ("Anyon superconductor", lambda th: H.mcsh_params(th, 42, 36), lambda th, pi, pf: H.mcsh_field(th, pi, pf, seed=5)),
    ("Ferromagnetic superconductor", lambda th: H.fs_params(th, 40, 34, eta2=0.05),
     lambda th, pi, pf: H.fs_field(th, pi, pf, seed=6)),
    ("Spin triplet superconducting ferromagnet", lambda th: H.st_params(th, 40, 34),
     lambda th, pi, pf: H.st_field(th, pi, pf, seed=6)),
    ("Liquid crystal", lambda th: H.lc_params(th, 40, 34), lambda th, pi, pf: H.lc_field(th, pi, pf, seed=6)),
]


@pytest.mark.parametrize("name,mkp,mkf", CASES, ids=[c[0] for c in CASES])
def test_exact_adjoint_matches_torch_autograd_of_energy(name, mkp, mkf):
    th = ssa.load_theory(name)
    p = mkp(th)
    p_i, p_f = H.packed(th, p)
    F = mkf(th, p_i, p_f)
    E_t, G_t = _torch_grad(H.ORACLES[name], F, p_i, p_f)
    sim = ssa.Simulation(p, th)
    sim.initialize({"mode": "array", "field": F})
    Fd = torch.tensor(F, device="cuda")
    G = sag.exact_gradient(sim, Fd).cpu().numpy()
    E = float(sag.energy_of(sim, Fd))
    scale = np.abs(G_t).max()
    print(f"[{name}] |E-E_t|/|E|={abs(E - E_t) / abs(E_t):.2e} max|G-G_t|/max|G_t|={np.abs(G - G_t).max() / scale:.2e}")
    assert abs(E - E_t) <= 1e-12 * max(1.0, abs(E_t))
    assert np.abs(G - G_t).max() <= 1e-11 * scale
    G2 = sim.energy_gradient("exact_adjoint")
    assert np.abs(G2 - G_t).max() <= 1e-11 * scale
    sim.close()


def test_gradcheck_on_patch_of_512_anisotropic():
    th = ssa.load_theory("Anisotropic superconductor")
    p = H.as_params(th, 512, 512, seed=0)
    p_i, p_f = H.packed(th, p)
    F = H.as_field(th, p_i, p_f, seed=0, amp=1e-2)
    sim = ssa.Simulation(p, th)
    base = torch.tensor(F, device="cuda")
    x0, y0 = 240, 250
    patch = base[:, x0:x0 + 32, y0:y0 + 32].clone().requires_grad_(True)

    def f(pt):
        full = base.clone()
        full[:, x0:x0 + 32, y0:y0 + 32] = pt
        return sag.discrete_energy(full, sim)

    ok = torch.autograd.gradcheck(f, (patch,), eps=1e-6, atol=1e-6, rtol=1e-4, nondet_tol=0.0, fast_mode=False)
    assert ok
    sim.close()


def _as_relax_case():
    """128x128 s+id vortex start in a 25x25 box, betas from seed 1, time step 0.3 h:
    the anisotropic kinetic term (coefficient 2 gamma) puts the stiffest mode of
    the exact-adjoint flow near the RK4 limit at the default 0.5 h."""
    th = ssa.load_theory("Anisotropic superconductor")
    n, xs = 128, 25.0
    p = H.as_params(th, n, n, xsize=xs, seed=1, time_step=0.3 * xs / (n - 1))
    p_i, p_f = H.packed(th, p)
    return th, p, p_i, p_f, H.as_field(th, p_i, p_f, seed=1, amp=1e-2)


def test_exact_flow_reaches_stationary_point_of_discrete_energy():
    th, p, p_i, p_f, F = _as_relax_case()
    sim = ssa.Simulation(p, th, variation="exact_adjoint")
    sim.initialize({"mode": "array", "field": F})
    E0 = sim.compute_energy()
    E, ea, n = sim.minimize(tol=1e-7, max_steps=120000, check_every=1000)
    G = sag.exact_gradient(sim, torch.tensor(sim.get_field(), device="cuda")).cpu().numpy()
    print(f"[exact flow] E0={E0:.10f} E={E:.12f} steps={n} max|EG|={ea:.2e} max|dE/dphi|/gv={np.abs(G).max() / p_f[4]:.2e}")
    assert E < E0
    assert ea < 1e-7
    assert np.abs(G).max() / p_f[4] < 1e-7
    sim.close()


@pytest.fixture(scope="module")
def relaxed_anisotropic():
    th, p, p_i, p_f, F = _as_relax_case()
    sim = ssa.Simulation(p, th, variation="exact_adjoint")
    sim.initialize({"mode": "array", "field": F})
    E, ea, n = sim.minimize(tol=1e-7, max_steps=120000, check_every=1000)
    print(f"[relaxed anisotropic 128x128] E*={E:.12f} max|EG|={ea:.2e} steps={n}", flush=True)
    assert ea < 1e-7
    Fstar = sim.get_field()
    grads = sag.energy_parameter_gradient(sim, torch.tensor(Fstar, device="cuda"),
                                          ["beta1", "beta2", "beta3", "beta4"])
    sim.close()
    return th, p, Fstar, grads


@pytest.mark.parametrize("name", ["beta1", "beta2", "beta3", "beta4"])
def test_parameter_sensitivity_matches_rerelaxation(relaxed_anisotropic, name):
    """dE*/dbeta_i via the envelope theorem vs central differences of
    re-relaxed energies, (E*(b+d) - E*(b-d))/(2d), at d and d/2; their
    Richardson combination (4 D(d/2) - D(d))/3 cancels the O(d^2) truncation."""
    th, p, Fstar, grads = relaxed_anisotropic
    g = grads[name]
    v = getattr(p, name)
    d = 1e-3 * max(abs(v), 1.0)
    Es = {}
    for dd in (d, d / 2):
        for s in (+1, -1):
            q = p.with_(**{name: v + s * dd})
            s2 = ssa.Simulation(q, th, variation="exact_adjoint")
            s2.initialize({"mode": "array", "field": Fstar})
            E, ea2, n2 = s2.minimize(tol=1e-7, max_steps=60000, check_every=1000)
            print(f"[re-relax {name}{'+' if s > 0 else '-'}{dd:.1e}] E*={E:.12f} max|EG|={ea2:.2e} steps={n2}",
                  flush=True)
            # a slow mode of this lattice relaxes below ~1e-7 only over >1e5 steps
            assert ea2 < 1e-6
            Es[(dd, s)] = E
            s2.close()
    fd = (Es[(d, 1)] - Es[(d, -1)]) / (2 * d)
    fd_half = (Es[(d / 2, 1)] - Es[(d / 2, -1)]) / d
    rich = (4.0 * fd_half - fd) / 3.0
    print(f"[dE*/d{name}] envelope={g:.10e} rerelax={fd:.10e} rel={abs(g - fd) / abs(fd):.2e} "
          f"richardson={rich:.10e} rel={abs(g - rich) / abs(rich):.2e}")
    # measured: 1.7e-6 - 6.6e-6 against the central difference at d (its own
    # O(d^2) truncation), 1.1e-9 - 6.3e-9 against the Richardson combination
    assert abs(g - fd) <= 2e-5 * abs(fd)
    assert abs(g - rich) <= 1e-7 * abs(rich)


_THETA = ["alpha1", "alpha2", "beta1", "beta2", "beta3", "beta4", "gamma1", "gamma2", "gamma12", "kappa", "Q"]


def test_parameter_gradient_by_backward_matches_energy_differences():
    """dE_h/dtheta from the backward of DiscreteEnergy (dual numbers in the
    packed couplings, chained through ResolvedParams) against the 5-point
    central difference in theta of the GPU energy at the same fixed field, for
    every coupling of the anisotropic theory (u1, u2 depend on alpha and beta,
    theta_diff and beta3tilde on beta4)."""
    th, p, p_i, p_f, F = _as_relax_case()
    sim = ssa.Simulation(p, th)
    try:
        field = torch.tensor(F, device="cuda")
        theta = torch.tensor([float(getattr(p, n)) for n in _THETA], dtype=torch.float64, device="cuda",
                             requires_grad=True)
        E = sag.discrete_energy(field, sim, theta, _THETA)
        E.backward()
        g = theta.grad.cpu().numpy()
        for k, name in enumerate(_THETA):
            v = float(getattr(p, name))
            h = 1e-3 * max(abs(v), 0.1)
            vals = []
            for j in (-2, -1, 1, 2):
                sim.set_params(p.with_(**{name: v + j * h}))
                vals.append(float(sag.energy_of(sim, field)))
            sim.set_params(p)
            fd = (vals[0] - 8.0 * vals[1] + 8.0 * vals[2] - vals[3]) / (12.0 * h)
            print(f"[dE/d{name}] backward {g[k]:.12e} 5-point {fd:.12e} rel {abs(g[k] - fd) / max(abs(fd), 1e-30):.1e}",
                  flush=True)
            assert abs(g[k] - fd) <= 1e-8 * max(abs(fd), 1e-3 * abs(float(E.detach())))
    finally:
        sim.close()


def test_parameter_gradient_needs_a_supported_theory():
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 40, 36)
    sim = ssa.Simulation(p, th)
    try:
        field = torch.tensor(sim.get_field(), device="cuda")
        with pytest.raises(ValueError, match="not available"):
            sag.packed_gradient(sim, field)
    finally:
        sim.close()

