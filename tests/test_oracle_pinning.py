"""Pins the oracle against what mathematics fixes (the reference ships no
tests, fixtures or golden vectors for this path).

Each check is chosen so that a plausible transcription mistake -- a wrong
stencil coefficient or sign, a transposed derivative index, a dropped term in
the variation, a wrong RK4 weight -- makes it fail.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import core as ocore
from oracle.theories import chiral_magnet as o_cm
from oracle.theories import ginzburg_landau as o_gl

import soliton_solver_amd as ssa
from tests import helpers as H


def _pp(X=24, Y=20, lsx=0.3, lsy=0.3, halo=2):
    p_i = np.array([X, Y, halo, 2, 1, X * Y, X * Y, 1, 1, 0], dtype=np.int32)
    p_f = np.array([lsx * (X - 1), lsy * (Y - 1), lsx, lsy, lsx * lsy, 0.5 * lsx])
    return p_i, p_f


def _poly_field(p_i, p_f, coeffs):
    X, Y = int(p_i[0]), int(p_i[1])
    x = p_f[2] * np.arange(X)[:, None] + 0.37
    y = p_f[3] * np.arange(Y)[None, :] - 0.21
    f = np.zeros((X, Y))
    for (a, b), c in coeffs.items():
        f = f + c * x ** a * y ** b
    return f[None], x, y


@pytest.mark.parametrize("lsx,lsy", [(0.3, 0.3), (0.25, 0.4)])
def test_first_derivative_exact_to_degree_4_and_zero_band(lsx, lsy):
    p_i, p_f = _pp(lsx=lsx, lsy=lsy)
    c = {(4, 0): 0.3, (3, 1): -0.2, (0, 4): 0.1, (2, 2): 0.05, (1, 0): 1.5, (0, 1): -0.7}
    F, x, y = _poly_field(p_i, p_f, c)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    fx = 4 * 0.3 * x ** 3 + 3 * -0.2 * x ** 2 * y + 2 * 0.05 * x * y ** 2 + 1.5
    fy = -0.2 * x ** 3 + 4 * 0.1 * y ** 3 + 2 * 0.05 * x ** 2 * y - 0.7
    h = 2
    X, Y = F.shape[1:]
    assert np.allclose(D1[0, 0, h:X - h, :], fx[h:X - h, :] * np.ones((1, Y)), rtol=1e-9, atol=1e-9)
    assert np.allclose(D1[0, 1, :, h:Y - h], (fy * np.ones((X, 1)))[:, h:Y - h], rtol=1e-9, atol=1e-9)
    assert np.all(D1[0, 0, :h] == 0) and np.all(D1[0, 0, X - h:] == 0)
    assert np.all(D1[0, 1, :, :h] == 0) and np.all(D1[0, 1, :, Y - h:] == 0)


@pytest.mark.parametrize("lsx,lsy", [(0.3, 0.3), (0.25, 0.4)])
def test_second_derivatives_exact_and_mixed_formula(lsx, lsy):
    """xx, yy exact to degree 5; the mixed entry equals
    f_xy + (lsx/lsy - 1) f_xx/2 + (lsy/lsx - 1) f_yy/2 (exact for polynomials
    of degree <= 5 along the diagonal) and is 0 unless halo < x, y (strict)."""
    p_i, p_f = _pp(lsx=lsx, lsy=lsy)
    F, x, y = _poly_field(p_i, p_f, {(5, 0): 0.01, (0, 5): -0.02, (2, 2): 0.3, (3, 1): 0.1, (1, 1): 2.0})
    D2 = ocore.second_derivatives(F, p_i, p_f)
    fxx = 20 * 0.01 * x ** 3 + 2 * 0.3 * y ** 2 + 6 * 0.1 * x * y
    fyy = 20 * -0.02 * y ** 3 + 2 * 0.3 * x ** 2
    fxy = 4 * 0.3 * x * y + 3 * 0.1 * x ** 2 + 2.0
    X, Y = F.shape[1:]
    h = 2
    fxx, fyy, fxy = (np.broadcast_to(a, (X, Y)) for a in (fxx, fyy, fxy))
    assert np.allclose(D2[0, 0, h:X - h], fxx[h:X - h], rtol=1e-8, atol=1e-8)
    assert np.allclose(D2[0, 1, :, h:Y - h], fyy[:, h:Y - h], rtol=1e-8, atol=1e-8)
    mixed = fxy + (lsx / lsy - 1) * fxx / 2 + (lsy / lsx - 1) * fyy / 2
    assert np.allclose(D2[0, 2, h + 1:X - h, h + 1:Y - h], mixed[h + 1:X - h, h + 1:Y - h], rtol=1e-8, atol=1e-8)
    assert np.all(D2[0, 2, h, :] == 0) and np.all(D2[0, 2, :, h] == 0)   # strict lower bound
    assert np.all(D2[0, 2, X - h:, :] == 0) and np.all(D2[0, 2, :, Y - h:] == 0)


def test_reference_sum_matches_fsum_and_tree_order():
    rng = np.random.default_rng(0)
    for n in (1, 1023, 1024, 1025, 5000):
        v = rng.normal(size=n) * 10.0 ** rng.integers(-3, 3, size=n)
        assert abs(ocore.reference_sum(v) - math.fsum(v)) <= 1e-13 * np.abs(v).sum()
    # tree order: pairs (i, i+512) are added first inside a block
    v = np.zeros(1024)
    v[0], v[512] = 1e16, 1.0
    v[1] = -1e16
    assert ocore.reference_sum(v) == (1e16 + 1.0) + (-1e16)


def test_reference_max_is_signed():
    assert ocore.reference_max(np.array([-5.0, -3.0, -4.0])) == -3.0
    assert ocore.reference_max(np.array([])) == -1.0e300


class _Linear:
    """Test theory with EG = w^2 phi pointwise (no stencil)."""
    HAS_CONSTRAINT = False
    W2 = 3.0

    @classmethod
    def energy_density(cls, F, D1, p_i, p_f):
        return 0.5 * cls.W2 * (F * F).sum(0) * p_f[4]

    @classmethod
    def variation(cls, F, D1, D2, p_i, p_f):
        return cls.W2 * F


def test_rk4_step_is_taylor_polynomial_of_exponential():
    """One step of phi'' = -w^2 phi equals sum_{k<=4} (dt A)^k/k! applied to (phi, v)."""
    p_i, p_f = _pp(X=8, Y=8)
    p_i[7] = 0  # no kill
    dt = p_f[5]
    rng = np.random.default_rng(1)
    F = rng.normal(size=(1, 8, 8))
    V = rng.normal(size=(1, 8, 8))
    F0, V0 = F.copy(), V.copy()
    ocore.arrested_newton_flow_step(_Linear, F, V, p_i, p_f, np.inf)
    A = np.array([[0.0, 1.0], [-_Linear.W2, 0.0]])
    M = np.eye(2)
    term = np.eye(2)
    for k in range(1, 5):
        term = term @ (dt * A) / k
        M = M + term
    assert np.allclose(F, M[0, 0] * F0 + M[0, 1] * V0, rtol=1e-14, atol=1e-14)
    assert np.allclose(V, M[1, 0] * F0 + M[1, 1] * V0, rtol=1e-14, atol=1e-14)


def test_kill_and_no_kill():
    p_i, p_f = _pp(X=8, Y=8)
    rng = np.random.default_rng(2)
    F = rng.normal(size=(1, 8, 8))
    V = rng.normal(size=(1, 8, 8))
    ocore.arrested_newton_flow_step(_Linear, F.copy(), V, p_i, p_f, -np.inf)
    assert np.all(V == 0)
    V = rng.normal(size=(1, 8, 8))
    ocore.arrested_newton_flow_step(_Linear, F.copy(), V, p_i, p_f, np.inf)
    assert np.abs(V).max() > 0
    p_i[7] = 0
    ocore.arrested_newton_flow_step(_Linear, F.copy(), V, p_i, p_f, -np.inf)
    assert np.abs(V).max() > 0


def test_kill_is_strict():
    """The arrest compares with a strict ``>`` (core/integrator.py:341): a step
    whose new energy equals the caller's prev_energy to the bit keeps its
    velocity.  The equal energy is produced by running the same step twice from
    the same state, the second time with prev_energy set to the first result."""
    p_i, p_f = _pp(X=8, Y=8)
    rng = np.random.default_rng(5)
    F0 = rng.normal(size=(1, 8, 8))
    V0 = rng.normal(size=(1, 8, 8))
    V = V0.copy()
    E1, _ = ocore.arrested_newton_flow_step(_Linear, F0.copy(), V, p_i, p_f, np.inf)
    assert np.abs(V).max() > 0
    V = V0.copy()
    E2, _ = ocore.arrested_newton_flow_step(_Linear, F0.copy(), V, p_i, p_f, E1)
    assert E2 == E1
    assert np.abs(V).max() > 0
    V = V0.copy()
    ocore.arrested_newton_flow_step(_Linear, F0.copy(), V, p_i, p_f, np.nextafter(E1, -np.inf))
    assert np.all(V == 0)


# --- variation against the energy: directional derivatives ------------------

def _bump(p_i, p_f, cx=0.5, cy=0.5, w=0.18):
    X, Y = int(p_i[0]), int(p_i[1])
    x = np.linspace(0, 1, X)[:, None]
    y = np.linspace(0, 1, Y)[None, :]
    r2 = ((x - cx) ** 2 + (y - cy) ** 2) / w ** 2
    return np.where(r2 < 1, np.exp(-1.0 / np.maximum(1 - r2, 1e-300)), 0.0)


def _dir_mismatch(o, F, Psi, p_i, p_f, eg_scale=1.0, comps=None):
    eps = 1e-5
    Ep = ocore.theory_energy(o, F + eps * Psi, p_i, p_f)
    Em = ocore.theory_energy(o, F - eps * Psi, p_i, p_f)
    dE = (Ep - Em) / (2 * eps)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    EG = o.variation(F.copy(), D1, D2, p_i, p_f)
    sel = comps if comps is not None else range(F.shape[0])
    G = sum((EG[a] * Psi[a]).sum() for a in sel) * p_f[4] * eg_scale
    return abs(G - dE) / abs(dE)


def test_chiral_variation_is_functional_derivative_of_energy():
    """sum gv EG.psi = dE_h(phi + e psi)/de + O(h^4) for smooth psi tangent to m
    (psi component: EG_psi = -coup_mu dE/dpsi, the reference's saddle form)."""
    th = ssa.load_theory("Chiral magnet")
    errs = []
    for n in (61, 121):
        p = H.chiral_params(th, n, n, xsize=4.0, demag=True, B=0.2)  # B != 0: the Zeeman term counts
        p_i, p_f = H.packed(th, p)
        g = H.grid_of(p_i, p_f)
        F = np.zeros((4, n, n))
        F[0] = 0.6 * np.sin(1.3 * g[0] + 0.4 * g[1])
        F[1] = 0.5 * np.cos(0.7 * g[1] - 0.9 * g[0])
        F[2] = 1.0
        F[:3] /= np.sqrt((F[:3] ** 2).sum(0))
        b = _bump(p_i, p_f, w=0.35)
        F[3] = 0.3 * b
        w = np.stack([b * np.cos(3.0 * np.linspace(0, 1, n))[:, None] * (1 + 0 * b), b * 0.5, b * 0.2])
        w = w - (w * F[:3]).sum(0) * F[:3]  # tangent to m
        Psi = np.zeros_like(F)
        Psi[:3] = w
        e_m = _dir_mismatch(o_cm, F, Psi, p_i, p_f, comps=range(3))
        Psi2 = np.zeros_like(F)
        Psi2[3] = b * (1 + 0.5 * np.sin(np.linspace(0, 3, n))[None, :])
        e_psi = _dir_mismatch(o_cm, F, Psi2, p_i, p_f, eg_scale=-1.0 / p_f[8], comps=[3])
        errs.append((e_m, e_psi))
    print("chiral directional mismatch (h, h/2):", errs)
    for k in range(2):
        assert errs[1][k] < 1e-3
        assert errs[0][k] / errs[1][k] > 8.0   # ~16x for a 4th-order scheme


def test_gl_variation_is_functional_derivative_of_energy():
    th = ssa.load_theory("Abelian Higgs")
    errs = []
    for n in (61, 121):
        lsx = 8.0 / (n - 1)
        p = H.gl_params(th, n, n, lsx=lsx, Lambda=1.3, q=0.8, u1=1.0, vortex_number=1.0)
        p_i, p_f = H.packed(th, p)
        F = np.zeros((4, n, n))
        V = np.zeros_like(F)
        th.initial_config.initialize(Field=F, Velocity=V, grid=H.grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                     config={"mode": "vortex", "scale": 1.0})
        b = _bump(p_i, p_f, w=0.3)
        F[0] += 0.2 * b
        F[1] -= 0.1 * b
        # a direction off the vortex centre: centred, psi.Psi integrates to zero
        # over the angle and the potential term (lambda) would drop out
        c = _bump(p_i, p_f, cx=0.6, cy=0.42, w=0.25)
        Psi = np.stack([c, 0.5 * c, -0.3 * c, 0.7 * c])
        errs.append(_dir_mismatch(o_gl, F, Psi, p_i, p_f))
    print("GL directional mismatch (h, h/2):", errs)
    assert errs[1] < 5e-4
    assert errs[0] / errs[1] > 8.0


def test_gl_energy_reads_vortex_number_not_u1():
    """Quirk kept from the reference: energy uses p_f[9], variation p_f[8]."""
    th = ssa.load_theory("Abelian Higgs")
    p = H.gl_params(th, 24, 24, u1=1.0, vortex_number=2.0)
    p_i, p_f = H.packed(th, p)
    F = H.gl_field(th, p_i, p_f)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    e2 = o_gl.energy_density(F, D1, p_i, p_f)
    p_f1 = p_f.copy()
    p_f1[8] = 7.0  # u1 must not matter for the energy
    assert np.array_equal(e2, o_gl.energy_density(F, D1, p_i, p_f1))


def test_skyrmion_number_of_ansatz():
    """Q of the reference ansatz: Bloch and Neel -1, anti +1 under the
    reference's density m.(d_x m x d_y m)/(4 pi)."""
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 161, 161, xsize=10.0)
    p_i, p_f = H.packed(th, p)
    for kind, sign in (("bloch", -1), ("neel", -1), ("anti", 1)):
        F = np.zeros((4, 161, 161))
        th.initial_config.initialize(Field=F, Velocity=np.zeros_like(F), grid=H.grid_of(p_i, p_f), p_i_h=p_i,
                                     p_f_h=p_f, config={"mode": "initial", "ansatz": kind})
        D1 = ocore.first_derivatives(F, p_i, p_f)
        Q = ocore.reference_sum(o_cm.skyrmion_density(F, D1, p_i, p_f))
        assert abs(Q - sign) < 2e-3, (kind, Q)


def test_unit_norm_and_projection_invariants_after_step():
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 32, 36)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=1, rough=True)
    V = np.zeros_like(F)
    ocore.relax(o_cm, F, V, p_i, p_f, 3)
    assert np.allclose((F[:3] ** 2).sum(0), 1.0, atol=1e-14)
    assert np.abs((V[:3] * F[:3]).sum(0)).max() < 1e-14
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    EG = o_cm.variation(F, D1, D2, p_i, p_f)
    assert np.abs((EG[:3] * F[:3]).sum(0)).max() < 1e-10 * np.abs(EG).max()
    assert np.all(EG[:, ocore.band_mask(p_i)] == 0)


def test_ground_state_is_stationary():
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 24, 24)
    p_i, p_f = H.packed(th, p)
    F = np.zeros((4, 24, 24))
    F[2] = 1.0
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    # the reference's stencil order leaves ~1e-16 round-off on a constant field
    assert np.abs(o_cm.variation(F, D1, D2, p_i, p_f)).max() < 1e-14
    assert abs(ocore.theory_energy(o_cm, F, p_i, p_f)) < 1e-14


_BS_POTENTIALS = ["standard", "holomorphic", "easy plane", "dihedral2", "aloof", "dihedral3", "broken",
                  "double vacua"]


@pytest.mark.parametrize("pot", _BS_POTENTIALS)
def test_baby_skyrme_variation_is_functional_derivative(pot):
    """Skyrme-term Euler-Lagrange expression and every potential derivative agree
    with dE_h along smooth tangent directions, converging at 4th order."""
    from oracle.theories import baby_skyrme as o_bs
    th = ssa.load_theory("Baby Skyrme model")
    errs = []
    for n in (61, 121):
        p = H.bs_params(th, n, n, xsize=4.0, kappa=0.9, mpi=0.8, potential=pot, N=2)
        p_i, p_f = H.packed(th, p)
        g = H.grid_of(p_i, p_f)
        F = np.zeros((3, n, n))
        F[0] = 0.6 * np.sin(1.3 * g[0] + 0.4 * g[1])
        F[1] = 0.5 * np.cos(0.7 * g[1] - 0.9 * g[0])
        F[2] = 0.4 + 0.2 * np.sin(g[0])
        F /= np.sqrt((F ** 2).sum(0))
        b = _bump(p_i, p_f, w=0.35)
        w = np.stack([b * np.cos(1.5 * g[1]), b * 0.5, b * (0.2 + 0.1 * g[0])])
        w = w - (w * F).sum(0) * F
        errs.append(_dir_mismatch(o_bs, F, w, p_i, p_f))
    print(pot, "baby Skyrme directional mismatch (h, h/2):", errs)
    assert errs[1] < 1e-3
    assert errs[0] / errs[1] > 8.0


def _as_smooth(n, lsx):
    x = lsx * np.arange(n)[:, None]
    y = lsx * np.arange(n)[None, :]
    F = np.zeros((6, n, n))
    F[0] = 0.3 * np.sin(0.7 * x + 0.2 * y)
    F[1] = 0.2 * np.cos(0.5 * y - 0.3 * x)
    F[2] = 0.6 + 0.2 * np.sin(x)
    F[3] = 0.3 * np.cos(y)
    F[4] = 0.5 * np.sin(0.4 * x + 0.3 * y)
    F[5] = 0.4 + 0.1 * np.cos(x * y / 5)
    return F, x


@pytest.mark.parametrize("part", ["gauge", "kinetic"])
def test_anisotropic_variation_matches_energy_without_potential(part):
    """Gauge components: EG = dE_h/dA at 4th order.  Order-parameter components with
    the potential switched off: EG = dE_h/dpsi at 4th order."""
    from oracle.theories import anisotropic_superconductor as o_as
    th = ssa.load_theory("Anisotropic superconductor")
    errs = []
    for n in (61, 121):
        lsx = 6.0 / (n - 1)
        kw = dict(xsize=6.0)
        if part == "kinetic":
            kw.update(alpha1=0.0, alpha2=0.0, beta1=0.0, beta2=0.0, beta3=0.0, beta4=0.0)
        p = H.as_params(th, n, n, **kw)
        p_i, p_f = H.packed(th, p)
        F, x = _as_smooth(n, lsx)
        b = _bump(p_i, p_f, w=0.35)
        Psi = np.zeros_like(F)
        comps = [0, 1] if part == "gauge" else [2, 3, 4, 5]
        for c in comps:
            Psi[c] = b * (1 + 0.3 * c * np.sin(x + c))
        errs.append(_dir_mismatch(o_as, F, Psi, p_i, p_f))
    print(part, "anisotropic directional mismatch (h, h/2):", errs)
    assert errs[1] < 1e-3
    assert errs[0] / errs[1] > 8.0


def test_anisotropic_potential_term_is_twice_its_derivative():
    """Reference quirk kept: with the anisotropy tensor switched off (no kinetic
    term for psi) the order-parameter variation is exactly 2 dV/dpsi."""
    from oracle.theories import anisotropic_superconductor as o_as
    th = ssa.load_theory("Anisotropic superconductor")
    n = 41
    p = H.as_params(th, n, n, xsize=6.0, gamma1=0.0, gamma2=0.0, gamma12=0.0, seed=3)
    p_i, p_f = H.packed(th, p)
    F, x = _as_smooth(n, 6.0 / (n - 1))
    b = _bump(p_i, p_f, w=0.35)
    Psi = np.zeros_like(F)
    Psi[2:] = b
    eps = 1e-5
    dE = (ocore.theory_energy(o_as, F + eps * Psi, p_i, p_f) - ocore.theory_energy(o_as, F - eps * Psi, p_i, p_f)) / (2 * eps)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    EG = o_as.variation(F, D1, D2, p_i, p_f)
    G = (EG * Psi).sum() * p_f[4]
    assert abs(G - 2.0 * dE) < 1e-8 * abs(dE)


def test_anisotropic_vacuum_has_zero_energy():
    """The subtracted constant makes the uniform vacuum psi1 = u1, psi2 = u2 e^{i theta} energy-free."""
    from oracle.theories import anisotropic_superconductor as o_as
    th = ssa.load_theory("Anisotropic superconductor")
    p = H.as_params(th, 24, 24, seed=1)
    p_i, p_f = H.packed(th, p)
    F = H.as_field(th, p_i, p_f, amp=0.0, mode="ground")
    E = ocore.theory_energy(o_as, F, p_i, p_f)
    assert abs(E) < 1e-12 * 24 * 24


def test_bec_variation_is_functional_derivative():
    from oracle.theories import bose_einstein_condensate as o_bec
    th = ssa.load_theory("BEC")
    errs = []
    for n in (61, 121):
        p = H.bec_params(th, n, n, xsize=6.0, beta=3.0, omega_rot=0.4)
        p_i, p_f = H.packed(th, p)
        g = H.grid_of(p_i, p_f)
        # a winding field and a direction with angular structure, so that the
        # rotation term -Omega L_z carries a good part of the derivative
        xp, yp = g[0] - 3.0, g[1] - 3.0
        f = np.exp(-(xp ** 2 + yp ** 2) / 3.0)
        F = np.stack([0.6 * f * (xp + 0.2), 0.6 * f * (yp - 0.1)])
        b = _bump(p_i, p_f, w=0.35)
        Psi = np.stack([b * (xp - 0.5 * yp), b * (yp + 0.3 * xp)])
        errs.append(_dir_mismatch(o_bec, F, Psi, p_i, p_f))
    print("BEC directional mismatch (h, h/2):", errs)
    assert errs[1] < 1e-3
    assert errs[0] / errs[1] > 8.0


def test_bec_step_keeps_unit_norm():
    from oracle.theories import bose_einstein_condensate as o_bec
    th = ssa.load_theory("BEC")
    p = H.bec_params(th, 40, 40)
    p_i, p_f = H.packed(th, p)
    F = H.bec_field(th, p_i, p_f)
    V = np.zeros_like(F)
    ocore.relax(o_bec, F, V, p_i, p_f, 3)
    assert abs(((F ** 2).sum(0) * p_f[4]).sum() - 1.0) < 1e-13


def test_mcsh_variation_matches_energy_and_a0_quirk():
    """With kappa = 0: EG = dE_h/dphi at 4th order for A0 = 0; with A0 != 0 the
    psi equations carry -A0^2 where dE/dpsi has +A0^2 (reference quirk), so
    EG.Psi - dE = sum gv (-2 q^2 A0^2 psi).Psi, also at 4th order."""
    from oracle.theories import maxwell_chern_simons_higgs as o_m
    th = ssa.load_theory("Anyon superconductor")
    for with_a0 in (False, True):
        errs = []
        for n in (61, 121):
            p = H.mcsh_params(th, n, n, xsize=8.0, kappa=0.0, q=0.9)
            p_i, p_f = H.packed(th, p)
            g = H.grid_of(p_i, p_f)
            F = np.zeros((5, n, n))
            F[0] = 0.3 * np.sin(0.7 * g[0] + 0.2 * g[1])
            F[1] = 0.2 * np.cos(0.5 * g[1] - 0.3 * g[0])
            F[2] = 0.7 + 0.2 * np.sin(0.5 * g[0])
            F[3] = 0.3 * np.cos(0.4 * g[1])
            F[4] = (0.25 * np.exp(-((g[0] - 4) ** 2 + (g[1] - 4) ** 2) / 6.0)) if with_a0 else 0.0
            b = _bump(p_i, p_f, w=0.35)
            Psi = np.stack([b, 0.5 * b, 0.8 * b, -0.4 * b, 0.6 * b])
            eps = 1e-5
            dE = (ocore.theory_energy(o_m, F + eps * Psi, p_i, p_f) - ocore.theory_energy(o_m, F - eps * Psi, p_i, p_f)) / (2 * eps)
            D1 = ocore.first_derivatives(F, p_i, p_f)
            D2 = ocore.second_derivatives(F, p_i, p_f)
            EG = o_m.variation(F, D1, D2, p_i, p_f)
            G = (EG * Psi).sum() * p_f[4]
            q = float(p_f[6])
            quirk = ((-2.0 * q * q * F[4] ** 2) * (F[2] * Psi[2] + F[3] * Psi[3])).sum() * p_f[4]
            errs.append(abs(G - dE - quirk) / abs(dE))
        print("MCSH mismatch (A0 on=%s):" % with_a0, errs)
        assert errs[1] < 1e-3
        assert errs[0] / errs[1] > 8.0


def _fs_smooth(th, n, **kw):
    p = H.fs_params(th, n, n, xsize=8.0, **kw)
    p_i, p_f = H.packed(th, p)
    g = H.grid_of(p_i, p_f)
    M0 = float(p_f[17])
    F = np.zeros((8, n, n))
    F[0] = 0.6 * np.sin(0.5 * g[0] + 0.2 * g[1])
    F[1] = 0.5 * np.cos(0.3 * g[1] - 0.4 * g[0])
    F[2] = 0.8 + 0.2 * np.sin(0.4 * g[0])
    F[:3] *= M0 / np.sqrt((F[:3] ** 2).sum(0))
    F[3] = 0.3 * np.sin(0.7 * g[0] + 0.2 * g[1])
    F[4] = 0.2 * np.cos(0.5 * g[1] - 0.3 * g[0])
    F[5] = 0.25 * np.sin(0.3 * g[0] * g[1] / 8.0)
    F[6] = 1.5 + 0.3 * np.sin(0.5 * g[0])
    F[7] = 0.4 * np.cos(0.4 * g[1])
    b = _bump(p_i, p_f, w=0.35)
    w = np.stack([b * np.cos(0.6 * g[1]), 0.5 * b, b * (0.2 + 0.1 * g[0]), 0.7 * b, -0.4 * b, 0.9 * b, 0.5 * b,
                  -0.8 * b])
    w[:3] -= (w[:3] * F[:3]).sum(0) * F[:3] / M0 ** 2
    return p_i, p_f, F, w


@pytest.mark.parametrize("with_eta2", [False, True])
def test_ferromagnetic_sc_variation_matches_energy_and_eta2_quirks(with_eta2):
    """With eta2 = 0: sum gv EG.w = dE_h/de at 4th order for w tangent to m (|m| = M0).
    With eta2 != 0 the reference's variation (kernels.py:591-614) drops the
    grad|psi|^2 . grad m_a term of d/dm (eta2 |grad m|^2 |psi|^2) and uses
    psi_b^2 where d/dpsi_b has psi_b, so
    EG.w - dE = sum gv 2 eta2 [grad|psi|^2 . grad m . w_m + |grad m|^2 (psi_b^2 - psi_b) w_psi_b]."""
    from oracle.theories import ferromagnetic_superconductor as o_fs
    th = ssa.load_theory("Ferromagnetic superconductor")
    errs = []
    for n in (61, 121):
        p_i, p_f, F, w = _fs_smooth(th, n, eta2=(0.3 if with_eta2 else 0.0))
        eps = 1e-5
        dE = (ocore.theory_energy(o_fs, F + eps * w, p_i, p_f) - ocore.theory_energy(o_fs, F - eps * w, p_i, p_f)) / (2 * eps)
        D1 = ocore.first_derivatives(F, p_i, p_f)
        D2 = ocore.second_derivatives(F, p_i, p_f)
        EG = o_fs.variation(F.copy(), D1, D2, p_i, p_f)
        G = (EG * w).sum() * p_f[4]
        e2 = float(p_f[10])
        psq_d = [2 * (F[6] * D1[6, i] + F[7] * D1[7, i]) for i in (0, 1)]
        gradm2 = sum(D1[a, i] ** 2 for a in range(3) for i in (0, 1))
        quirk = sum(2 * e2 * (psq_d[0] * D1[a, 0] + psq_d[1] * D1[a, 1]) * w[a] for a in range(3))
        quirk = quirk + sum(2 * e2 * gradm2 * (F[b] ** 2 - F[b]) * w[b] for b in (6, 7))
        errs.append(abs(G - dE - quirk.sum() * p_f[4]) / abs(dE))
        if with_eta2:
            assert abs(quirk.sum() * p_f[4]) > 1e-2 * abs(dE)
    print("ferromagnetic SC mismatch (eta2 on=%s):" % with_eta2, errs)
    assert errs[1] < 1e-3
    assert errs[0] / errs[1] > 8.0


def test_ferromagnetic_sc_vacuum_and_constraint():
    """The coupled vacuum m = (0, 0, M0), psi = u1 has zero energy (the subtracted
    constant), and one step keeps |m| = M0 with V tangent to m."""
    from oracle.theories import ferromagnetic_superconductor as o_fs
    th = ssa.load_theory("Ferromagnetic superconductor")
    p = H.fs_params(th, 24, 24)
    p_i, p_f = H.packed(th, p)
    F = np.zeros((8, 24, 24))
    V = np.zeros_like(F)
    th.initial_config.initialize(Field=F, Velocity=V, grid=H.grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                 config={"mode": "ground"})
    assert abs(ocore.theory_energy(o_fs, F, p_i, p_f)) < 1e-12 * 24 * 24
    F = H.fs_field(th, p_i, p_f, seed=2, amp=0.05)
    V = np.zeros_like(F)
    ocore.relax(o_fs, F, V, p_i, p_f, 2)
    M0 = float(p_f[17])
    assert np.abs(np.sqrt((F[:3] ** 2).sum(0)) - M0).max() < 1e-13
    assert np.abs((V[:3] * F[:3]).sum(0)).max() < 1e-12


def test_ferromagnetic_sc_ansatz_charges():
    """The skyrmion-vortex ansatz carries Q = -1 for the Bloch texture of winding 1
    (m3 = M0 cos f, f(0) = pi), and no flux through the x or y faces (A3 = 0)."""
    from oracle.theories import ferromagnetic_superconductor as o_fs
    th = ssa.load_theory("Ferromagnetic superconductor")
    p = H.fs_params(th, 161, 161, xsize=40.0)
    p_i, p_f = H.packed(th, p)
    F = H.fs_field(th, p_i, p_f, amp=0.0)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    Q = ocore.reference_sum(o_fs.skyrmion_density(F, D1, p_i, p_f).ravel())
    assert abs(Q + 1.0) < 2e-3
    for which in (0, 1):
        assert ocore.reference_sum(o_fs.vortex_density(F, D1, p_i, p_f, which).ravel()) == 0.0


def test_spin_triplet_variation_is_functional_derivative():
    """sum gv EG.w = dE_h/de at 4th order for w tangent to m (|m| = M0), every
    field perturbed, both order parameters, the hb2 and hc couplings switched on."""
    from oracle.theories import spin_triplet as o_st
    th = ssa.load_theory("Spin triplet superconducting ferromagnet")
    errs = []
    for n in (61, 121):
        p = H.st_params(th, n, n, xsize=8.0, hb2=0.7, hc=-0.4)
        p_i, p_f = H.packed(th, p)
        g = H.grid_of(p_i, p_f)
        M0 = float(p_f[20])
        F = np.zeros((10, n, n))
        F[0] = 0.6 * np.sin(0.5 * g[0] + 0.2 * g[1])
        F[1] = 0.5 * np.cos(0.3 * g[1] - 0.4 * g[0])
        F[2] = 0.8 + 0.2 * np.sin(0.4 * g[0])
        F[:3] *= M0 / np.sqrt((F[:3] ** 2).sum(0))
        F[3] = 0.3 * np.sin(0.7 * g[0] + 0.2 * g[1])
        F[4] = 0.2 * np.cos(0.5 * g[1] - 0.3 * g[0])
        F[5] = 0.25 * np.sin(0.3 * g[0] * g[1] / 8.0)
        F[6] = 1.1 + 0.3 * np.sin(0.5 * g[0])
        F[7] = 0.4 * np.cos(0.4 * g[1])
        F[8] = 0.9 + 0.2 * np.cos(0.3 * g[0] - 0.2 * g[1])
        F[9] = -0.3 * np.sin(0.6 * g[1])
        b = _bump(p_i, p_f, w=0.35)
        w = np.stack([b * np.cos(0.6 * g[1]), 0.5 * b, b * (0.2 + 0.1 * g[0]), 0.7 * b, -0.4 * b, 0.9 * b, 0.5 * b,
                      -0.8 * b, 0.6 * b * np.sin(g[0]), 0.3 * b])
        w[:3] -= (w[:3] * F[:3]).sum(0) * F[:3] / M0 ** 2
        errs.append(_dir_mismatch(o_st, F, w, p_i, p_f))
    print("spin-triplet directional mismatch (h, h/2):", errs)
    assert errs[1] < 1e-3
    assert errs[0] / errs[1] > 8.0


def test_spin_triplet_vacuum_and_ansatz_charge():
    from oracle.theories import spin_triplet as o_st
    th = ssa.load_theory("Spin triplet superconducting ferromagnet")
    p = H.st_params(th, 24, 24)
    p_i, p_f = H.packed(th, p)
    F = np.zeros((10, 24, 24))
    th.initial_config.initialize(Field=F, Velocity=np.zeros_like(F), grid=H.grid_of(p_i, p_f), p_i_h=p_i,
                                 p_f_h=p_f, config={"mode": "ground"})
    assert abs(ocore.theory_energy(o_st, F, p_i, p_f)) < 1e-12 * 24 * 24
    p = H.st_params(th, 161, 161, xsize=40.0, ansatz="neel")
    p_i, p_f = H.packed(th, p)
    F = H.st_field(th, p_i, p_f, amp=0.0, ansatz="neel")
    D1 = ocore.first_derivatives(F, p_i, p_f)
    assert abs(ocore.reference_sum(o_st.skyrmion_density(F, D1, p_i, p_f).ravel()) + 1.0) < 2e-3


def _lc_smooth(th, n, unit=True, **kw):
    """A smooth director (unit, or -- ``unit`` False -- of length 1 +- 0.15, as
    the stage fields of a step are, which are not renormalised) with the
    depolarisation potential Phi, and a direction w tangent to it."""
    p = H.lc_params(th, n, n, xsize=8.0, **kw)
    p_i, p_f = H.packed(th, p)
    g = H.grid_of(p_i, p_f)
    F = np.zeros((4, n, n))
    F[0] = 0.6 * np.sin(0.5 * g[0] + 0.2 * g[1])
    F[1] = 0.5 * np.cos(0.3 * g[1] - 0.4 * g[0])
    F[2] = 0.8 + 0.2 * np.sin(0.4 * g[0])
    F[:3] /= np.sqrt((F[:3] ** 2).sum(0))
    if not unit:
        F[:3] *= 1.0 + 0.15 * np.cos(0.35 * g[0] - 0.25 * g[1])
    F[3] = 0.4 * np.sin(0.3 * g[0] + 0.5 * g[1])
    b = _bump(p_i, p_f, w=0.35)
    w = np.zeros_like(F)
    w[:3] = np.stack([b * np.cos(0.6 * g[1]), 0.5 * b, b * (0.2 + 0.1 * g[0])])
    w[:3] -= (w[:3] * F[:3]).sum(0) * F[:3] / (F[:3] ** 2).sum(0)
    return p_i, p_f, F, w


@pytest.mark.parametrize("deformation", ["twist", "splay-bend"])
def test_liquid_crystal_variation_is_functional_derivative_without_depol(deformation):
    from oracle.theories import liquid_crystal as o_lc
    th = ssa.load_theory("Liquid crystal")
    errs = []
    for n in (61, 121):
        p_i, p_f, F, w = _lc_smooth(th, n, deformation=deformation)
        errs.append(_dir_mismatch(o_lc, F, w, p_i, p_f))
    print(deformation, "liquid crystal directional mismatch (h, h/2):", errs)
    assert errs[1] < 1e-3
    assert errs[0] / errs[1] > 8.0


@pytest.mark.parametrize("unit", [True, False])
def test_liquid_crystal_depolarisation_director_terms_are_twice_their_derivative(unit):
    """Reference quirk kept: with depol on, the director variation's
    depolarisation part (kernels.py:492-500) is 2 d/dn of the depolarisation
    energy (kernels.py:203-209), so EG.w - dE = dE_depol at 4th order.  Also on
    a director that is not of unit length, where the (e3/e1) n.(d_i n) d_i Phi
    term, zero for unit directors, contributes."""
    from oracle.theories import liquid_crystal as o_lc
    th = ssa.load_theory("Liquid crystal")
    errs = []
    eps = 1e-5
    for n in (61, 121):
        d = {}
        for depol in (False, True):
            p_i, p_f, F, w = _lc_smooth(th, n, unit=unit, depol=depol)
            d[depol] = (ocore.theory_energy(o_lc, F + eps * w, p_i, p_f) - ocore.theory_energy(o_lc, F - eps * w, p_i, p_f)) / (2 * eps)
        D1 = ocore.first_derivatives(F, p_i, p_f)
        D2 = ocore.second_derivatives(F, p_i, p_f)
        G = (o_lc.variation(F.copy(), D1, D2, p_i, p_f) * w).sum() * p_f[4]
        dE_depol = d[True] - d[False]
        assert abs(dE_depol) > 0.1 * abs(d[True])
        errs.append(abs(G - d[True] - dE_depol) / abs(d[True]))
    print("liquid crystal depol quirk mismatch (h, h/2):", errs)
    assert errs[1] < 1e-3
    assert errs[0] / errs[1] > 8.0


def test_liquid_crystal_phi_equation_is_poisson_with_bound_charge():
    """EG_Phi = -lap Phi + rho/coup_eps with rho = div P,
    P_i = n_i div n - (e3/e1)(n.grad) n_i (in-plane i): checked against the
    stencil divergence of P (4th order) and, exactly, against the electric
    charge density of the observables (rho = -density/gv)."""
    from oracle.theories import liquid_crystal as o_lc
    th = ssa.load_theory("Liquid crystal")
    errs = []
    for n in (61, 121):
        p_i, p_f, F, _ = _lc_smooth(th, n, depol=True)
        D1 = ocore.first_derivatives(F, p_i, p_f)
        D2 = ocore.second_derivatives(F, p_i, p_f)
        EG = o_lc.variation(F.copy(), D1, D2, p_i, p_f)
        gv, ceps, r = float(p_f[4]), float(p_f[8]), float(p_f[10]) / float(p_f[9])
        rho = -o_lc.electric_charge_density(F, D1, D2, p_i, p_f) / gv
        inner = ~ocore.band_mask(p_i)
        lhs = EG[3] + D2[3, 0] + D2[3, 1]
        assert np.abs((lhs - rho / ceps)[inner]).max() < 1e-12 * max(1.0, np.abs(rho).max())
        divn = D1[0, 0] + D1[1, 1]
        P = np.stack([F[i] * divn - r * (F[0] * D1[i, 0] + F[1] * D1[i, 1]) for i in (0, 1)])
        DP = ocore.first_derivatives(np.concatenate([P, P]), p_i, p_f)
        div_P = DP[0, 0] + DP[1, 1]
        core = np.zeros_like(inner)
        core[6:-6, 6:-6] = True
        errs.append(np.abs((div_P - rho)[core]).max() / np.abs(rho[core]).max())
    print("liquid crystal div P vs rho (h, h/2):", errs)
    assert errs[1] < 1e-4
    assert errs[0] / errs[1] > 8.0


@pytest.mark.parametrize("unit", [True, False])
def test_liquid_crystal_depolarisation_energy_is_flexoelectric_polarisation(unit):
    """The depolarisation energy density (kernels.py:203-209) equals
    (gv/2) grad Phi . (n div n + (e3/e1) n x curl n), the flexoelectric
    polarisation (splay e1, bend e3) with d_z = 0.  The identity
    (n.grad) n = grad(|n|^2)/2 - n x curl n is algebraic in the stencil
    derivatives, so it holds to rounding, also for a director whose length
    varies, where the (e3/e1) n.(d_i n) d_i Phi / 2 term does not vanish."""
    from oracle.theories import liquid_crystal as o_lc
    th = ssa.load_theory("Liquid crystal")
    p_i1, p_f1, F, _ = _lc_smooth(th, 61, unit=unit, depol=True)
    p_i0, p_f0, _, _ = _lc_smooth(th, 61, unit=unit, depol=False)
    D1 = ocore.first_derivatives(F, p_i1, p_f1)
    e_depol = o_lc.energy_density(F, D1, p_i1, p_f1) - o_lc.energy_density(F, D1, p_i0, p_f0)
    gv, r = float(p_f1[4]), float(p_f1[10]) / float(p_f1[9])
    assert r != 1.0
    divn = D1[0, 0] + D1[1, 1]
    curl = np.stack([D1[2, 1], -D1[2, 0], D1[1, 0] - D1[0, 1]])
    n_x_curl = np.cross(F[:3], curl, axis=0)
    P = F[:2] * divn + r * n_x_curl[:2]
    want = 0.5 * gv * (D1[3, 0] * P[0] + D1[3, 1] * P[1])
    scale = np.abs(want).max()
    assert np.abs(e_depol - want).max() < 1e-12 * scale
    if not unit:   # the |n|-gradient part is a visible share of the density
        grad_n2 = np.stack([(F[:3] * D1[:3, i]).sum(0) for i in (0, 1)])
        assert np.abs(0.25 * r * gv * (grad_n2 * D1[3, :2]).sum(0)).max() > 0.05 * scale


def test_aux_oracle_supercurrent_is_curl_of_flux_and_phase_difference_is_normalised_dot():
    """Supercurrent (d_y B3, -d_x B3, -lap A3) with B3 = d_x A2 - d_y A1, at 4th order
    in the interior; cos of the phase difference = psi1.psi2/(|psi1||psi2|);
    the Higgs density of a field built in polar form rho e^{i phi} is rho^2."""
    from oracle import aux as oaux
    th = ssa.load_theory("Ferromagnetic superconductor")
    errs = []
    for n in (61, 121):
        p_i, p_f, F, _ = _fs_smooth(th, n)
        D1 = ocore.first_derivatives(F, p_i, p_f)
        D2 = ocore.second_derivatives(F, p_i, p_f)
        J = oaux.supercurrent(D2, 3, 4, 5)
        B = D1[4, 0] - D1[3, 1]
        Bd = ocore.first_derivatives(np.stack([B] * 8), p_i, p_f)[0]
        ref = np.stack([Bd[1], -Bd[0]])
        c = slice(8, -8)
        errs.append(np.abs(J[:2, c, c] - ref[:, c, c]).max() / np.abs(ref[:, c, c]).max())
        lap = D2[5, 0] + D2[5, 1]
        assert np.array_equal(J[2], -lap)
    print("supercurrent vs curl B (h, h/2):", errs)
    assert errs[1] < 1e-4 and errs[0] / errs[1] > 8.0
    rng = np.random.default_rng(0)
    G = rng.normal(size=(6, 9, 9))
    cosd = oaux.phase_difference(G)
    dot = (G[2] * G[4] + G[3] * G[5]) / np.sqrt((G[2] ** 2 + G[3] ** 2) * (G[4] ** 2 + G[5] ** 2))
    assert np.abs(cosd - dot).max() < 1e-12
    # |psi|^2 of psi = rho e^{i phi} built in polar form is rho^2
    rho = rng.uniform(0.1, 2.0, size=(9, 9))
    phi = rng.uniform(-np.pi, np.pi, size=(9, 9))
    H2 = np.stack([np.zeros((9, 9)), np.zeros((9, 9)), rho * np.cos(phi), rho * np.sin(phi)])
    assert np.abs(oaux.higgs(H2, 2, 3) - rho ** 2).max() < 1e-14 * 4.0
