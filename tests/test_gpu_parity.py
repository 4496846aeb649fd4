"""HIP path vs the FP64 oracle on identical seeded inputs, every theory.

Tolerances (FP64 form): per-stage energy gradient within 1e-11 of max|EG|;
energies and charges within 1e-12 relative for a given field; after N steps
energy within 1e-10 relative and max |Delta F| stated (asserted <= 1e-9).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import soliton_solver_amd as ssa
from soliton_solver_amd import _lib
from oracle import core as ocore
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _sim(theory, params, F, V=None, precision="fp64"):
    sim = ssa.Simulation(params, theory, precision=precision)
    sim.initialize({"mode": "array", "field": F, "velocity": V})
    return sim


@pytest.mark.parametrize("shape", [(48, 40), (100, 150), (37, 70)])
@pytest.mark.parametrize("rough", [False, True])
def test_chiral_variation_energy_charges(shape, rough):
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, shape[0], shape[1])
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=3, rough=rough)
    sim = _sim(th, p, F)
    o = H.ORACLES["Chiral magnet"]
    eg_gpu = sim.energy_gradient()
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    eg_ref = o.variation(F.copy(), D1, D2, p_i, p_f)
    scale = np.abs(eg_ref).max()
    print(f"[chiral {shape} rough={rough}] max|EG|={scale:.3e} maxdiff={np.abs(eg_gpu - eg_ref).max():.3e}")
    assert np.abs(eg_gpu - eg_ref).max() <= 1e-11 * scale
    E_ref = ocore.reference_sum(o.energy_density(F, D1, p_i, p_f).ravel())
    assert H.rel(sim.compute_energy(), E_ref) < 1e-12
    obs = sim.observables()
    Q_ref = ocore.reference_sum(o.skyrmion_density(F, D1, p_i, p_f).ravel())
    M_ref = ocore.reference_sum(o.magnetic_charge_density(F, D1, p_i, p_f).ravel())
    assert abs(obs["skyrmion_number"] - Q_ref) < 1e-11
    assert abs(obs["magnetic_charge"] - M_ref) < 1e-11 * max(1.0, np.abs(o.magnetic_charge_density(F, D1, p_i, p_f)).sum())
    dens = sim.compute_energy_density()
    assert np.allclose(dens, o.energy_density(F, D1, p_i, p_f), rtol=1e-12, atol=1e-14)
    sim.close()


@pytest.mark.parametrize("shape", [(48, 40), (66, 130)])
def test_gl_variation_energy_charge(shape):
    th = ssa.load_theory("Abelian Higgs")
    p = H.gl_params(th, shape[0], shape[1])
    p_i, p_f = H.packed(th, p)
    F = H.gl_field(th, p_i, p_f, seed=5)
    sim = _sim(th, p, F)
    o = H.ORACLES["Ginzburg-Landau superconductor"]
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    eg_ref = o.variation(F, D1, D2, p_i, p_f)
    eg_gpu = sim.energy_gradient()
    scale = np.abs(eg_ref).max()
    assert np.abs(eg_gpu - eg_ref).max() <= 1e-11 * scale
    E_ref = ocore.reference_sum(o.energy_density(F, D1, p_i, p_f).ravel())
    assert H.rel(sim.compute_energy(), E_ref) < 1e-12
    N_ref = ocore.reference_sum(o.vortex_density(F, D1, p_i, p_f).ravel())
    assert abs(sim.observables()["vortex_number"] - N_ref) < 1e-11
    sim.close()


def _trajectory(name, th, p, F, V, nsteps, chunks):
    p_i, p_f = H.packed(th, p)
    o = H.ORACLES[name]
    Fo, Vo = F.copy(), V.copy()
    E0 = ocore.theory_energy(o, Fo, p_i, p_f)
    Es, errs = ocore.relax(o, Fo, Vo, p_i, p_f, nsteps, prev_energy=E0)
    sim = _sim(th, p, F, V)
    E0g = sim.compute_energy()
    assert H.rel(E0g, E0) < 1e-12
    E = E0g
    gE, gerr = [], []
    for n in chunks:
        E, err, _ = sim.steps(E, n, history=True)
        gE.extend(sim.last_history[0])
        gerr.extend(sim.last_history[1])
    Fg, Vg = sim.get_field(), sim.get_velocity()
    sim.close()
    return np.array(Es), np.array(errs), Fo, Vo, np.array(gE), np.array(gerr), Fg, Vg


@pytest.mark.parametrize("chunks", [[1] * 12, [5, 7], [12]])
def test_chiral_steps_match_oracle(chunks):
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 40, 52)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=11)
    V = np.zeros_like(F)
    n = sum(chunks)
    Es, errs, Fo, Vo, gE, gerr, Fg, Vg = _trajectory("Chiral magnet", th, p, F, V, n, chunks)
    dF = np.abs(Fg - Fo).max()
    print(f"[chiral steps {chunks}] E={gE[-1]:.15e} relE={H.rel(gE[-1], Es[-1]):.2e} max|dF|={dF:.2e}")
    assert np.all(np.abs(gE - Es) <= 1e-10 * np.abs(Es))
    assert np.allclose(gerr, errs, rtol=1e-9, atol=1e-9 * np.abs(errs).max())
    assert dF <= 1e-9
    assert np.abs(Vg - Vo).max() <= 1e-9 * max(1.0, np.abs(Vo).max())


def test_gl_steps_match_oracle_with_velocity():
    th = ssa.load_theory("Abelian Higgs")
    p = H.gl_params(th, 48, 56)
    p_i, p_f = H.packed(th, p)
    F, V = H.gl_field(th, p_i, p_f, seed=2, velocity=True)
    Es, errs, Fo, Vo, gE, gerr, Fg, Vg = _trajectory("Ginzburg-Landau superconductor", th, p, F, V, 30, [10, 20])
    assert np.all(np.abs(gE - Es) <= 1e-10 * np.abs(Es))
    assert np.abs(Fg - Fo).max() <= 1e-9
    assert np.abs(Vg - Vo).max() <= 1e-9


def test_kill_semantics_forced_from_host():
    """prev_energy = -inf forces a kill every step; +inf never kills."""
    th = ssa.load_theory("Abelian Higgs")
    p = H.gl_params(th, 32, 32)
    p_i, p_f = H.packed(th, p)
    F, V = H.gl_field(th, p_i, p_f, seed=9, velocity=True)
    o = H.ORACLES["Ginzburg-Landau superconductor"]
    for prev in (-np.inf, np.inf):
        Fo, Vo = F.copy(), V.copy()
        Eo, erro = ocore.arrested_newton_flow_step(o, Fo, Vo, p_i, p_f, prev)
        sim = _sim(th, p, F, V)
        Eg, errg = sim.step(prev)
        Vg = sim.get_velocity()
        sim.close()
        assert H.rel(Eg, Eo) < 1e-12
        if prev == -np.inf:
            assert np.all(Vg == 0.0) and np.all(Vo == 0.0)
        else:
            assert np.abs(Vg - Vo).max() < 1e-12 and np.abs(Vo).max() > 0


def test_kill_is_strict_on_the_device():
    """The device arrest is a strict ``>`` (core/integrator.py:341): the same
    step from the same state, with prev_energy set to the energy that step
    produces (to the bit), keeps its velocity; with prev_energy one ulp below,
    it arrests."""
    th = ssa.load_theory("Abelian Higgs")
    p = H.gl_params(th, 32, 32)
    p_i, p_f = H.packed(th, p)
    F, V = H.gl_field(th, p_i, p_f, seed=11, velocity=True)
    sim = _sim(th, p, F, V)
    E1, _ = sim.step(np.inf)
    V1 = sim.get_velocity()
    assert np.abs(V1).max() > 0
    for prev, arrested in ((E1, False), (np.nextafter(E1, -np.inf), True)):
        sim.set_field(F)
        sim.set_velocity(V)
        E2, _ = sim.step(prev)
        V2 = sim.get_velocity()
        assert E2 == E1
        if arrested:
            assert np.all(V2 == 0.0)
        else:
            assert np.array_equal(V2, V1)
    sim.close()


def test_killkinen_off_never_kills():
    th = ssa.load_theory("Abelian Higgs")
    p = H.gl_params(th, 32, 32, killkinen=False)
    p_i, p_f = H.packed(th, p)
    F, V = H.gl_field(th, p_i, p_f, seed=9, velocity=True)
    sim = _sim(th, p, F, V)
    sim.step(-np.inf)
    assert np.abs(sim.get_velocity()).max() > 0
    sim.close()


@pytest.mark.parametrize("name", ["Anisotropic superconductor", "Bose-Einstein condensate"])
def test_graph_replay_of_unfused_flows_matches_eager(name):
    """The exact-adjoint flow (anisotropic) and the L2-normalised flow (BEC)
    allocate scratch on first use; with graphs on from the first call, that
    allocation happens before the capture and the replays equal eager steps."""
    th = ssa.load_theory(name)
    if name == "Bose-Einstein condensate":
        p = H.bec_params(th, 48, 40)
        p_i, p_f = H.packed(th, p)
        F, variation = H.bec_field(th, p_i, p_f, seed=2), "reference"
    else:
        p = H.as_params(th, 48, 40, seed=2, xsize=16.0)
        p_i, p_f = H.packed(th, p)
        F, variation = H.as_field(th, p_i, p_f, seed=3), "exact_adjoint"
    sims = []
    for graphs in (0, 4):
        s = ssa.Simulation(p, th, variation=variation)
        s.initialize({"mode": "array", "field": F})
        s.use_graphs(graphs)
        sims.append(s)
    a, b = sims
    Ea = Eb = a.compute_energy()
    for _ in range(3):
        Ea, erra, _ = a.steps(Ea, 10)
        Eb, errb, _ = b.steps(Eb, 10)
        assert Ea == Eb and erra == errb
    assert np.array_equal(a.get_field(), b.get_field())
    a.close()
    b.close()


def test_graph_replay_matches_eager():
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 64, 64)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=21)
    a = _sim(th, p, F)
    b = _sim(th, p, F)
    b.use_graphs(4)
    Ea = Eb = a.compute_energy()
    for _ in range(4):
        Ea, erra, _ = a.steps(Ea, 10)
        a.synchronize()
        Eb, errb, _ = b.steps(Eb, 10)
        b.synchronize()
        assert Ea == Eb and erra == errb
    assert np.array_equal(a.get_field(), b.get_field())
    assert np.array_equal(a.get_velocity(), b.get_velocity())
    a.close()
    b.close()


def test_fp32_form_tracks_fp64():
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 64, 64)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=4)
    a = _sim(th, p, F)
    b = _sim(th, p, F, precision="fp32")
    E0 = a.compute_energy()
    assert H.rel(b.compute_energy(), E0) < 1e-6
    Ea, _, _ = a.steps(E0, 50)
    Eb, _, _ = b.steps(E0, 50)
    print(f"[fp32 vs fp64] relE={H.rel(Eb, Ea):.2e} max|dm|={np.abs(a.get_field() - b.get_field()).max():.2e}")
    assert H.rel(Eb, Ea) < 1e-5
    # polish: switch to FP64 in place and keep relaxing
    b.set_precision("fp64")
    Eb2, _, _ = b.steps(Eb, 10)
    assert np.isfinite(Eb2)
    a.close()
    b.close()


_FP32_CASES = {
    "Ginzburg-Landau superconductor": (lambda th: H.gl_params(th, 60, 52), lambda th, pi, pf: H.gl_field(th, pi, pf, seed=3)),
    "Baby Skyrme model": (lambda th: H.bs_params(th, 60, 52), lambda th, pi, pf: H.bs_field(th, pi, pf, seed=3, k=2)),
    "Anisotropic superconductor": (lambda th: H.as_params(th, 60, 52, seed=2, xsize=16.0),
                                   lambda th, pi, pf: H.as_field(th, pi, pf, seed=3)),
    "Anyon superconductor": (lambda th: H.mcsh_params(th, 60, 52), lambda th, pi, pf: H.mcsh_field(th, pi, pf, seed=3)),
    "Ferromagnetic superconductor": (lambda th: H.fs_params(th, 60, 52, eta2=0.05),
                                     lambda th, pi, pf: H.fs_field(th, pi, pf, seed=3)),
    "Spin triplet superconducting ferromagnet": (lambda th: H.st_params(th, 60, 52),
                                                 lambda th, pi, pf: H.st_field(th, pi, pf, seed=3)),
    "Liquid crystal": (lambda th: H.lc_params(th, 60, 52), lambda th, pi, pf: H.lc_field(th, pi, pf, seed=3)),
}


@pytest.mark.parametrize("name", list(_FP32_CASES))
def test_fp32_form_tracks_fp64_for_every_theory(name):
    """FP32 fields and stencils with FP64 energies: the energy gradient agrees
    with the FP64 form to FP32 rounding (1e-5 of max|EG|), the energy to 1e-6,
    and 20 steps from the same field stay within 1e-5 relative energy."""
    th = ssa.load_theory(name)
    mkp, mkf = _FP32_CASES[name]
    p = mkp(th)
    p_i, p_f = H.packed(th, p)
    F = mkf(th, p_i, p_f)
    a = _sim(th, p, F)
    b = _sim(th, p, F, precision="fp32")
    ga, gb = a.energy_gradient(), b.energy_gradient()
    scale = np.abs(ga).max()
    E0 = a.compute_energy()
    print(f"[fp32 {name}] max|dEG|/max|EG|={np.abs(ga - gb).max() / scale:.2e} "
          f"relE0={H.rel(b.compute_energy(), E0):.2e}")
    assert np.abs(ga - gb).max() <= 1e-5 * scale
    assert H.rel(b.compute_energy(), E0) < 1e-6
    Ea, _, _ = a.steps(E0, 20)
    Eb, _, _ = b.steps(E0, 20)
    print(f"[fp32 {name}] after 20 steps relE={H.rel(Eb, Ea):.2e}")
    assert H.rel(Eb, Ea) < 1e-5
    a.close()
    b.close()


def test_invalid_arguments_raise():
    th = ssa.load_theory("Chiral magnet")
    with pytest.raises(ValueError):
        ssa.Simulation(H.chiral_params(th, 32, 32, number_total_fields=5), th)
    sim = ssa.Simulation(H.chiral_params(th, 32, 32), th)
    with pytest.raises(ValueError):
        sim.set_field(np.zeros(10))
    with pytest.raises(ValueError):
        sim.compute_charge(7)
    sim.close()


@pytest.mark.parametrize("pot", ["standard", "holomorphic", "easy plane", "dihedral2", "aloof", "dihedral3",
                                 "broken", "double vacua"])
def test_baby_skyrme_variation_energy_charge(pot):
    th = ssa.load_theory("Baby Skyrme model")
    p = H.bs_params(th, 70, 90, kappa=1.3, mpi=0.7, potential=pot, N=3)
    p_i, p_f = H.packed(th, p)
    F = H.bs_field(th, p_i, p_f, seed=4, k=2, amp=1e-2)
    sim = _sim(th, p, F)
    o = H.ORACLES["Baby Skyrme model"]
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    eg_ref = o.variation(F.copy(), D1, D2, p_i, p_f)
    eg_gpu = sim.energy_gradient()
    scale = np.abs(eg_ref).max()
    print(f"[bs {pot}] max|EG|={scale:.3e} maxdiff={np.abs(eg_gpu - eg_ref).max():.3e}")
    assert np.abs(eg_gpu - eg_ref).max() <= 1e-11 * scale
    E_ref = ocore.reference_sum(o.energy_density(F, D1, p_i, p_f).ravel())
    assert H.rel(sim.compute_energy(), E_ref) < 1e-12
    Q_ref = ocore.reference_sum(o.skyrmion_density(F, D1, p_i, p_f).ravel())
    assert abs(sim.observables()["skyrmion_number"] - Q_ref) < 1e-11
    sim.close()


def test_baby_skyrme_steps_match_oracle():
    th = ssa.load_theory("Baby Skyrme model")
    p = H.bs_params(th, 64, 64, kappa=1.0, mpi=0.5)
    p_i, p_f = H.packed(th, p)
    F = H.bs_field(th, p_i, p_f, seed=8, k=2)
    V = np.zeros_like(F)
    Es, errs, Fo, Vo, gE, gerr, Fg, Vg = _trajectory("Baby Skyrme model", th, p, F, V, 25, [25])
    assert np.all(np.abs(gE - Es) <= 1e-10 * np.abs(Es))
    assert np.abs(Fg - Fo).max() <= 1e-9


@pytest.mark.parametrize("seed", [0, 1])
def test_anisotropic_variation_energy_charge(seed):
    th = ssa.load_theory("Anisotropic superconductor")
    p = H.as_params(th, 72, 60, seed=seed)
    p_i, p_f = H.packed(th, p)
    F = H.as_field(th, p_i, p_f, seed=seed, amp=5e-2)
    sim = _sim(th, p, F)
    o = H.ORACLES["Anisotropic superconductor"]
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    eg_ref = o.variation(F, D1, D2, p_i, p_f)
    eg_gpu = sim.energy_gradient()
    scale = np.abs(eg_ref).max()
    print(f"[as seed={seed}] max|EG|={scale:.3e} maxdiff={np.abs(eg_gpu - eg_ref).max():.3e}")
    assert np.abs(eg_gpu - eg_ref).max() <= 1e-11 * scale
    E_ref = ocore.reference_sum(o.energy_density(F, D1, p_i, p_f).ravel())
    assert abs(sim.compute_energy() - E_ref) <= 1e-12 * np.abs(o.energy_density(F, D1, p_i, p_f)).sum()
    N_ref = ocore.reference_sum(o.vortex_density(F, D1, p_i, p_f).ravel())
    assert abs(sim.observables()["vortex_number"] - N_ref) < 1e-11
    sim.close()


def test_anisotropic_steps_match_oracle():
    th = ssa.load_theory("Anisotropic superconductor")
    p = H.as_params(th, 48, 52, seed=2)
    p_i, p_f = H.packed(th, p)
    F = H.as_field(th, p_i, p_f, seed=2)
    V = np.zeros_like(F)
    Es, errs, Fo, Vo, gE, gerr, Fg, Vg = _trajectory("Anisotropic superconductor", th, p, F, V, 20, [20])
    assert np.all(np.abs(gE - Es) <= 1e-10 * np.abs(Es))
    assert np.abs(Fg - Fo).max() <= 1e-9


def test_bec_variation_energy_and_normalised_steps():
    th = ssa.load_theory("Bose-Einstein condensate")
    p = H.bec_params(th, 56, 48)
    p_i, p_f = H.packed(th, p)
    F = H.bec_field(th, p_i, p_f, seed=3)
    o = H.ORACLES["Bose-Einstein condensate"]
    sim = _sim(th, p, F)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    eg_ref = o.variation(F, D1, D2, p_i, p_f)
    assert np.abs(sim.energy_gradient() - eg_ref).max() <= 1e-11 * np.abs(eg_ref).max()
    assert H.rel(sim.compute_energy(), ocore.reference_sum(o.energy_density(F, D1, p_i, p_f))) < 1e-12
    sim.close()
    V = np.zeros_like(F)
    Es, errs, Fo, Vo, gE, gerr, Fg, Vg = _trajectory("Bose-Einstein condensate", th, p, F, V, 20, [8, 12])
    assert np.all(np.abs(gE - Es) <= 1e-10 * np.abs(Es))
    assert np.abs(Fg - Fo).max() <= 1e-9
    assert abs((Fg ** 2).sum() * p_f[4] - 1.0) < 1e-12


def test_mcsh_variation_energy_charges_and_steps():
    th = ssa.load_theory("Anyon superconductor")
    p = H.mcsh_params(th, 60, 52)
    p_i, p_f = H.packed(th, p)
    F = H.mcsh_field(th, p_i, p_f, seed=2)
    o = H.ORACLES["Anyon superconductor"]
    sim = _sim(th, p, F)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    eg_ref = o.variation(F, D1, D2, p_i, p_f)
    assert np.abs(sim.energy_gradient() - eg_ref).max() <= 1e-11 * np.abs(eg_ref).max()
    assert H.rel(sim.compute_energy(), ocore.reference_sum(o.energy_density(F, D1, p_i, p_f))) < 1e-12
    obs = sim.observables()
    assert abs(obs["vortex_number"] - ocore.reference_sum(o.vortex_density(F, D1, p_i, p_f))) < 1e-11
    el = ocore.reference_sum(o.electric_charge_density(F, D1, D2, p_i, p_f))
    assert abs(obs["electric_charge"] - el) < 1e-10 * max(1.0, abs(el))
    nt = ocore.reference_sum(o.noether_charge_density(F, D1, p_i, p_f))
    assert abs(th.observables.compute_noether_charge(sim) - nt) < 1e-12 * max(1.0, abs(nt))
    sim.close()
    Es, errs, Fo, Vo, gE, gerr, Fg, Vg = _trajectory("Anyon superconductor", th, p, F, np.zeros_like(F), 15, [15])
    assert np.all(np.abs(gE - Es) <= 1e-10 * np.abs(Es))
    assert np.abs(Fg - Fo).max() <= 1e-9


@pytest.mark.parametrize("ansatz", ["bloch", "neel", "anti"])
def test_ferromagnetic_sc_variation_energy_charges_and_steps(ansatz):
    th = ssa.load_theory("Ferromagnetic superconductor")
    p = H.fs_params(th, 58, 50, eta2=0.05)
    p_i, p_f = H.packed(th, p)
    F = H.fs_field(th, p_i, p_f, seed=4, ansatz=ansatz)
    o = H.ORACLES["Ferromagnetic superconductor"]
    sim = _sim(th, p, F)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    eg_ref = o.variation(F.copy(), D1, D2, p_i, p_f)
    assert np.abs(sim.energy_gradient() - eg_ref).max() <= 1e-11 * np.abs(eg_ref).max()
    assert H.rel(sim.compute_energy(), ocore.reference_sum(o.energy_density(F, D1, p_i, p_f))) < 1e-12
    obs = sim.observables()
    assert abs(obs["skyrmion_number"] - ocore.reference_sum(o.skyrmion_density(F, D1, p_i, p_f))) < 1e-11
    for which in (2, 0, 1):
        ref = ocore.reference_sum(o.vortex_density(F, D1, p_i, p_f, which))
        assert abs(th.observables.compute_vortex_number(sim, which) - ref) < 1e-11
    sim.close()
    Es, errs, Fo, Vo, gE, gerr, Fg, Vg = _trajectory("Ferromagnetic superconductor", th, p, F, np.zeros_like(F), 12, [5, 7])
    print(f"[fs {ansatz}] relE={H.rel(gE[-1], Es[-1]):.2e} max|dF|={np.abs(Fg - Fo).max():.2e}")
    assert np.all(np.abs(gE - Es) <= 1e-10 * np.abs(Es))
    assert np.abs(Fg - Fo).max() <= 1e-9
    M0 = float(p_f[17])
    assert np.abs(np.sqrt((Fg[:3] ** 2).sum(0)) - M0).max() < 1e-12


def test_spin_triplet_variation_energy_charges_and_steps():
    name = "Spin triplet superconducting ferromagnet"
    th = ssa.load_theory(name)
    p = H.st_params(th, 54, 46, vortex2_number=2.0, hb2=0.7, hc=-0.4)
    p_i, p_f = H.packed(th, p)
    F = H.st_field(th, p_i, p_f, seed=6)
    o = H.ORACLES[name]
    sim = _sim(th, p, F)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    eg_ref = o.variation(F.copy(), D1, D2, p_i, p_f)
    assert np.abs(sim.energy_gradient() - eg_ref).max() <= 1e-11 * np.abs(eg_ref).max()
    assert H.rel(sim.compute_energy(), ocore.reference_sum(o.energy_density(F, D1, p_i, p_f))) < 1e-12
    obs = sim.observables()
    assert abs(obs["skyrmion_number"] - ocore.reference_sum(o.skyrmion_density(F, D1, p_i, p_f))) < 1e-11
    for which in (2, 0, 1):
        ref = ocore.reference_sum(o.vortex_density(F, D1, p_i, p_f, which))
        assert abs(th.observables.compute_vortex_number(sim, which) - ref) < 1e-11
    sim.close()
    Es, errs, Fo, Vo, gE, gerr, Fg, Vg = _trajectory(name, th, p, F, np.zeros_like(F), 10, [4, 6])
    print(f"[spin triplet] relE={H.rel(gE[-1], Es[-1]):.2e} max|dF|={np.abs(Fg - Fo).max():.2e}")
    assert np.all(np.abs(gE - Es) <= 1e-10 * np.abs(Es))
    assert np.abs(Fg - Fo).max() <= 1e-9


@pytest.mark.parametrize("deformation,depol", [("twist", False), ("twist", True), ("splay-bend", True)])
def test_liquid_crystal_variation_energy_charges_and_steps(deformation, depol):
    name = "Liquid crystal"
    th = ssa.load_theory(name)
    p = H.lc_params(th, 62, 50, deformation=deformation, depol=depol)
    p_i, p_f = H.packed(th, p)
    F = H.lc_field(th, p_i, p_f, seed=8, ansatz="neel" if deformation == "splay-bend" else "bloch")
    o = H.ORACLES[name]
    sim = _sim(th, p, F)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    eg_ref = o.variation(F.copy(), D1, D2, p_i, p_f)
    assert np.abs(sim.energy_gradient() - eg_ref).max() <= 1e-11 * np.abs(eg_ref).max()
    assert H.rel(sim.compute_energy(), ocore.reference_sum(o.energy_density(F, D1, p_i, p_f))) < 1e-12
    obs = sim.observables()
    assert abs(obs["skyrmion_number"] - ocore.reference_sum(o.skyrmion_density(F, D1, p_i, p_f))) < 1e-11
    el = o.electric_charge_density(F, D1, D2, p_i, p_f)
    assert abs(obs["electric_charge"] - ocore.reference_sum(el)) < 1e-11 * max(1.0, np.abs(el).sum())
    sim.close()
    Es, errs, Fo, Vo, gE, gerr, Fg, Vg = _trajectory(name, th, p, F, np.zeros_like(F), 12, [5, 7])
    print(f"[lc {deformation} depol={depol}] relE={H.rel(gE[-1], Es[-1]):.2e} max|dF|={np.abs(Fg - Fo).max():.2e}")
    assert np.all(np.abs(gE - Es) <= 1e-10 * np.abs(Es))
    assert np.abs(Fg - Fo).max() <= 1e-9


@pytest.mark.parametrize("variants,fused,fused41", [
    ([1, 1, 1, 1], False, False), ([2, 2, 2, 2], False, False), ([0, 2, 1, 2], False, False),
    ([3, 3, 3, 3], False, False), ([4, 4, 4, 4], False, False), ([5, 5, 5, 5], False, False),
    ([0, 3, 4, 5], False, False), ([0, 3, 0, 3], True, False), ([2, 5, 0, 4], True, False),
    ([0, 1, 2, 3], False, True), ([1, 3, 0, 5], True, True), ([2, 4, 4, 4], True, True),
    ([6, 7, 8, 6], False, False), ([0, 6, 7, 8], True, False), ([7, 8, 1, 6], True, True),
    ([3, 6, 0, 7], False, True), ([0, 1, 0, 3], True, False), ([2, 2, 2, 2], True, False),
    ([1, 2, 1, 6], True, True), ([2, 2, 2, 2], True, True), ([1, 0, 1, 1], False, True),
    ([0, 1, 0, 2], True, True)])
def test_tile_shape_variants_match_oracle(variants, fused, fused41):
    """Every launch-shape variant of the fused passes (tiles and streaming strips,
    lattice 70 x 130: partial strips and partial column blocks), with and without
    the fused 2+3 and 4+1 sweeps, gives the oracle's trajectory."""
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 70, 130)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=21, rough=True)
    o = H.ORACLES["Chiral magnet"]
    Fo, Vo = F.copy(), np.zeros_like(F)
    E0 = ocore.theory_energy(o, Fo, p_i, p_f)
    Es, _ = ocore.relax(o, Fo, Vo, p_i, p_f, 10, prev_energy=E0)
    sim = _sim(th, p, F)
    sim.set_tile_variants(variants)
    sim.set_fused_stages(fused, fused41)
    info = sim.tile_variants()
    assert info["variants"] == variants and info["fused23"] == fused and info["fused41"] == fused41
    E, _, _ = sim.steps(sim.compute_energy(), 10, history=True)
    gE = np.array(sim.last_history[0])
    dF = np.abs(sim.get_field() - Fo).max()
    print(f"[tile variants {variants} rows {info['rows']}] relE={H.rel(gE[-1], Es[-1]):.2e} max|dF|={dF:.2e}")
    assert np.all(np.abs(gE - np.array(Es)) <= 1e-10 * np.abs(np.array(Es)))
    assert dF <= 1e-9
    sim.close()


def test_autotune_leaves_state_unchanged_and_steps_agree():
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 200, 256)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=5)
    a = _sim(th, p, F, precision="fp32")
    b = _sim(th, p, F, precision="fp32")
    E0 = a.compute_energy()
    a.steps(E0, 3)
    b.steps(E0, 3)
    Fa, Va = a.get_field(), a.get_velocity()
    info = a.autotune_tiles(reps=3)
    print(f"[autotune 200x256 fp32] {info}")
    assert np.array_equal(a.get_field(), Fa) and np.array_equal(a.get_velocity(), Va)
    assert all(0 <= v < 9 for v in info["variants"])
    Ea, _, _ = a.steps(E0, 8)
    Eb, _, _ = b.steps(E0, 8)
    assert H.rel(Ea, Eb) < 1e-12
    assert np.abs(a.get_field() - b.get_field()).max() <= 1e-6
    a.close()
    b.close()


@pytest.mark.parametrize("name", ["Ferromagnetic superconductor", "Spin triplet superconducting ferromagnet",
                                  "Baby Skyrme model"])
def test_streaming_form_matches_oracle_for_wide_theories(name):
    """Streaming strips with one row per wave (8 and 10 FP64 fields) and the
    unit-norm baby Skyrme model: same trajectory as the oracle."""
    th = ssa.load_theory(name)
    if name.startswith("Ferro"):
        p = H.fs_params(th, 90, 70, eta2=0.05)
        mk = lambda pi, pf: H.fs_field(th, pi, pf, seed=3)
    elif name.startswith("Spin"):
        p = H.st_params(th, 90, 70)
        mk = lambda pi, pf: H.st_field(th, pi, pf, seed=3)
    else:
        p = H.bs_params(th, 90, 70, xsize=10.0)
        mk = lambda pi, pf: H.bs_field(th, pi, pf, seed=3, k=2)
    p_i, p_f = H.packed(th, p)
    F = mk(p_i, p_f)
    o = H.ORACLES[name]
    Fo, Vo = F.copy(), np.zeros_like(F)
    Es, _ = ocore.relax(o, Fo, Vo, p_i, p_f, 6, prev_energy=ocore.theory_energy(o, Fo, p_i, p_f))
    sim = _sim(th, p, F)
    sim.set_tile_variants([3, 4, 5, 3])
    if name.startswith("Baby"):
        sim.set_fused_stages(True, True)  # 3 fields: both FP64 halves fit one ring
    sim.steps(sim.compute_energy(), 6, history=True)
    gE = np.array(sim.last_history[0])
    dF = np.abs(sim.get_field() - Fo).max()
    print(f"[stream {name}] relE={H.rel(gE[-1], Es[-1]):.2e} max|dF|={dF:.2e}")
    assert np.all(np.abs(gE - np.array(Es)) <= 1e-10 * np.abs(np.array(Es)))
    assert dF <= 1e-9
    sim.close()


def _relax_calls(sim, E, counts):
    """Consecutive calls of ``counts`` steps; returns the energies and errs of every step."""
    hE, hR = [], []
    for n in counts:
        E, _, _ = sim.steps(E, n, history=True)
        if n:
            hE += list(sim.last_history[0])
            hR += list(sim.last_history[1])
    return E, np.array(hE), np.array(hR)


@pytest.mark.parametrize("precision,graphs", [("fp64", 0), ("fp64", 4), ("fp32", 0), ("fp32", 6)])
def test_fused41_matches_separate_passes(precision, graphs):
    """The 4+1 schedule (pass 1 only in the first step of a call, l1 in two
    buffers) against the separate passes over calls of 0, 1, 2, 7, 13 and 1
    steps, eager and in graphs: the same energies, errs and fields (FP64 energy
    sums differ only in order)."""
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 150, 200)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=8, rough=True)
    a = _sim(th, p, F, precision=precision)
    b = _sim(th, p, F, precision=precision)
    a.set_tile_variants([3, 3, 3, 4])
    b.set_tile_variants([3, 3, 3, 4])
    a.set_fused_stages(True, True)
    b.set_fused_stages(True, False)
    if graphs:
        a.use_graphs(graphs)
        b.use_graphs(graphs)
    counts = [0, 1, 2, 7, 13, 1]
    E0 = a.compute_energy()
    Ea, hEa, hRa = _relax_calls(a, E0, counts)
    Eb, hEb, hRb = _relax_calls(b, E0, counts)
    dF = np.abs(a.get_field() - b.get_field()).max()
    dV = np.abs(a.get_velocity() - b.get_velocity()).max()
    print(f"[fused41 {precision} graphs={graphs}] relE={H.rel(Ea, Eb):.2e} max|dF|={dF:.2e} max|dV|={dV:.2e} "
          f"max hist relE={np.max(np.abs(hEa - hEb) / np.abs(hEb)):.2e}")
    tE, tF, tR = (1e-12, 1e-12, 1e-10) if precision == "fp64" else (1e-7, 1e-5, 1e-4)
    assert len(hEa) == sum(counts)
    assert np.all(np.abs(hEa - hEb) <= tE * np.abs(hEb))
    assert np.all(np.abs(hRa - hRb) <= tR * np.abs(hRb).max())
    assert H.rel(Ea, Eb) < tE
    assert dF <= tF and dV <= tF
    a.close()
    b.close()


@pytest.mark.parametrize("graphs", [0, 4])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_folded_finalize_is_bit_exact(precision, graphs):
    """On small lattices the 4+1 schedule folds each step's finalize (err, the
    new field's energy, the arrest decision) into the next step's 2+3 sweep,
    where every workgroup reduces the partials in the finalize kernel's order;
    the timed schedule (ss_time_passes) launches the finalize on its own.  After
    1, 2 and 9 steps from a rough field (arrests included) the fields and
    velocities agree bit for bit, and the returned energy is the new field's."""
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 96, 80)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=8, rough=True)
    for n in (1, 2, 9):
        a = _sim(th, p, F, precision=precision)
        b = _sim(th, p, F, precision=precision)
        for s in (a, b):
            s.set_fused_stages(True, True)
        if graphs:
            a.use_graphs(graphs)
        E0 = a.compute_energy()
        Ea, _, _ = a.steps(E0, n, history=True)
        hE = a.last_history[0]
        b.time_passes(E0, n)
        rises = int(np.sum(np.diff(np.concatenate([[E0], hE])) > 0))
        print(f"[folded finalize {precision} graphs={graphs} n={n}] energy rises={rises} "
              f"relE={H.rel(Ea, a.compute_energy()):.2e}")
        assert np.array_equal(a.get_field(), b.get_field())
        assert np.array_equal(a.get_velocity(), b.get_velocity())
        assert H.rel(Ea, a.compute_energy()) < (1e-12 if precision == "fp64" else 1e-6)
        a.close()
        b.close()


def _history(sim, n):
    """Per-step energies and errs of the context's last call, whatever ran it."""
    hE, hR = np.empty(n), np.empty(n)
    sim._check(_lib.lib().ss_step_history(sim._ctx, hE, hR, n))
    return hE, hR


@pytest.mark.parametrize("graphs", [0, 4])
@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
def test_deferred_err_finalize_is_bit_exact(precision, graphs):
    """The paired schedules finalize the err of a step together with the next
    step's energy (one launch after the next 1+2 sweep; the last step's err
    after the call); the timed schedule (ss_time_passes) finalizes every step on
    its own.  After 1, 2 and 9 steps from a rough field (arrests included) the
    fields and velocities agree bit for bit, so do the per-step energies and
    errs, and the call returns the last step's err."""
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 96, 80)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=8, rough=True)
    for n in (1, 2, 9):
        a = _sim(th, p, F, precision=precision)
        b = _sim(th, p, F, precision=precision)
        for s in (a, b):
            s.set_tile_variants([3, 3, 3, 3])
            s.set_fused_stages(False, False, pairs=True)
        assert a.tile_variants()["pairs"]
        if precision == "fp32c":
            assert a.tile_variants()["comp12"]
        if graphs:
            a.use_graphs(graphs)
        E0 = a.compute_energy()
        Ea, ema, _ = a.steps(E0, n, history=True)
        hEa, hRa = a.last_history
        b.time_passes(E0, n)
        hEb, hRb = _history(b, n)
        rises = int(np.sum(np.diff(np.concatenate([[E0], hEa])) > 0))
        print(f"[deferred err finalize {precision} graphs={graphs} n={n}] energy rises={rises}")
        assert np.array_equal(a.get_field(), b.get_field())
        assert np.array_equal(a.get_velocity(), b.get_velocity())
        assert np.array_equal(hEa, hEb) and np.array_equal(hRa, hRb)
        assert ema == hRa[-1]
        a.close()
        b.close()


def _paired_cases():
    return ["Chiral magnet", *_FP32_CASES]


@pytest.mark.parametrize("variants", [[3, 3, 3, 3], [6, 6, 6, 8]])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", _paired_cases())
def test_paired_schedule_matches_separate_passes(name, precision, variants):
    """Passes 1+2 and 3+4 as two sweeps (Temp2 = F + dt V/2 needs no l1; l3
    stays in registers for the combine) against the separate passes in the
    same strips: calls of 0, 1, 2, 7 and 5 steps (chiral: also in graphs of 4
    steps), then 2 steps from prev_energy = -inf, whose first step is arrested
    (passes 3+4 take l2 = l1): the same energies, errs, fields and velocities."""
    th = ssa.load_theory(name)
    if name == "Chiral magnet":
        p = H.chiral_params(th, 70, 130)
        p_i, p_f = H.packed(th, p)
        F = H.chiral_field(th, p_i, p_f, seed=5, rough=True)
    else:
        mkp, mkf = _FP32_CASES[name]
        p = mkp(th)
        p_i, p_f = H.packed(th, p)
        F = mkf(th, p_i, p_f)
    tE, tF, tR = (1e-12, 1e-12, 1e-10) if precision == "fp64" else (1e-7, 1e-5, 1e-4)
    a = _sim(th, p, F, precision=precision)
    b = _sim(th, p, F, precision=precision)
    a.set_tile_variants(variants)
    b.set_tile_variants(variants)
    try:
        a.set_fused_stages(False, False, pairs=True)
    except ValueError:  # the two-field ring does not fit this theory and precision
        a.close()
        b.close()
        pytest.skip(f"no paired sweeps for {name} in {precision}")
    assert a.tile_variants()["pairs"]
    if name == "Chiral magnet":
        a.use_graphs(4)
        b.use_graphs(4)
    E0 = a.compute_energy()
    counts = [0, 1, 2, 7, 5]
    Ea, hEa, hRa = _relax_calls(a, E0, counts)
    Eb, hEb, hRb = _relax_calls(b, E0, counts)
    scale = max(1.0, np.abs(F).max())
    dF = np.abs(a.get_field() - b.get_field()).max() / scale
    dV = np.abs(a.get_velocity() - b.get_velocity()).max() / scale
    print(f"[paired {name} {precision} {variants}] relE={H.rel(Ea, Eb):.2e} max|dF|={dF:.2e} max|dV|={dV:.2e}")
    assert len(hEa) == sum(counts)
    assert np.all(np.abs(hEa - hEb) <= tE * np.abs(hEb))
    assert np.all(np.abs(hRa - hRb) <= tR * np.abs(hRb).max())
    assert dF <= tF and dV <= tF
    # an arrested first step: V is zeroed and l2 = l1
    Ea, _, _ = a.steps(-np.inf, 2)
    Eb, _, _ = b.steps(-np.inf, 2)
    dF = np.abs(a.get_field() - b.get_field()).max() / scale
    dV = np.abs(a.get_velocity() - b.get_velocity()).max() / scale
    print(f"[paired {name} {precision} arrested] relE={H.rel(Ea, Eb):.2e} max|dF|={dF:.2e} max|dV|={dV:.2e}")
    assert H.rel(Ea, Eb) < tE
    assert dF <= tF and dV <= tF
    a.close()
    b.close()


def test_paired_schedule_refusals():
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 48, 40)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=1)
    s = _sim(th, p, F)
    with pytest.raises(ValueError):
        s.set_fused_stages(True, False, pairs=True)  # the pairs replace the other sweeps
    s.close()
    th = ssa.load_theory("Bose-Einstein condensate")
    s = ssa.Simulation(H.bec_params(th, 32, 32), th)
    with pytest.raises(ValueError):
        s.set_fused_stages(False, False, pairs=True)
    s.close()


_FUSED23_TILE_CASES = {
    "Chiral magnet": (lambda th: H.chiral_params(th, 70, 130),
                      lambda th, pi, pf: H.chiral_field(th, pi, pf, seed=5, rough=True)),
    **_FP32_CASES,
}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", list(_FUSED23_TILE_CASES))
def test_fused23_tiles_match_separate_passes(name, precision):
    """Passes 2+3 in one LDS-tile sweep (both stage fields in one tile) against
    the separate tile passes, in the tile variants whose two-field tile fits
    (the rest fall back to the streaming sweep, or have no fused sweep): the
    same fields and energies after 15 steps."""
    th = ssa.load_theory(name)
    mkp, mkf = _FUSED23_TILE_CASES[name]
    p = mkp(th)
    p_i, p_f = H.packed(th, p)
    F = mkf(th, p_i, p_f)
    tol = 1e-12 if precision == "fp64" else 1e-5
    ran_tile = False
    for v in (1, 2):
        a = _sim(th, p, F, precision=precision)
        b = _sim(th, p, F, precision=precision)
        a.set_tile_variants([v, v, v, v])
        b.set_tile_variants([v, v, v, v])
        try:
            b.set_fused_stages(True, False)
        except ValueError:  # neither a two-field tile nor a fused strip fits this theory and precision
            a.close()
            b.close()
            continue
        ran_tile |= b.tile_variants()["variants"][1] == v
        E0 = a.compute_energy()
        Ea, _, _ = a.steps(E0, 15)
        Eb, _, _ = b.steps(E0, 15)
        dF = np.abs(a.get_field() - b.get_field()).max() / max(1.0, np.abs(F).max())
        print(f"[fused23 tiles {name} {precision} v={v} -> {b.tile_variants()['variants']}] "
              f"relE={H.rel(Eb, Ea):.2e} max|dF|={dF:.2e}")
        assert H.rel(Eb, Ea) < tol and dF <= tol
        a.close()
        b.close()
    if name == "Chiral magnet":
        assert ran_tile  # FP64 and FP32 chiral tiles of 4..16 rows hold both fields


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", list(_FUSED23_TILE_CASES))
def test_fused41_tiles_match_separate_passes(name, precision):
    """Pass 4 + the next step's pass 1 in one LDS-tile sweep (Temp4 and F' in
    one tile, F' of the halo sites formed from their own loads) against the
    separate tile passes, over calls of 0, 1, 2, 7 and 5 steps (chiral: also in
    graphs of 4 steps): the same energies, errs and fields."""
    th = ssa.load_theory(name)
    mkp, mkf = _FUSED23_TILE_CASES[name]
    p = mkp(th)
    p_i, p_f = H.packed(th, p)
    F = mkf(th, p_i, p_f)
    tE, tF, tR = (1e-12, 1e-12, 1e-10) if precision == "fp64" else (1e-7, 1e-5, 1e-4)
    counts = [0, 1, 2, 7, 5]
    ran_tile = False
    for v, graphs in ((2, 0), (1, 0), (2, 4 if name == "Chiral magnet" else 0)):
        a = _sim(th, p, F, precision=precision)
        b = _sim(th, p, F, precision=precision)
        a.set_tile_variants([v, v, v, v])
        b.set_tile_variants([v, v, v, v])
        try:
            a.set_fused_stages(False, True)
        except ValueError:  # no 4+1 sweep for this theory and precision
            a.close()
            b.close()
            continue
        ran_tile |= a.tile_variants()["variants"][3] == v
        if graphs:
            a.use_graphs(graphs)
            b.use_graphs(graphs)
        E0 = a.compute_energy()
        Ea, hEa, hRa = _relax_calls(a, E0, counts)
        Eb, hEb, hRb = _relax_calls(b, E0, counts)
        scale = max(1.0, np.abs(F).max())
        dF = np.abs(a.get_field() - b.get_field()).max() / scale
        dV = np.abs(a.get_velocity() - b.get_velocity()).max() / scale
        print(f"[fused41 tiles {name} {precision} v={v} graphs={graphs} -> {a.tile_variants()['variants']}] "
              f"relE={H.rel(Ea, Eb):.2e} max|dF|={dF:.2e} max|dV|={dV:.2e}")
        assert len(hEa) == sum(counts)
        assert np.all(np.abs(hEa - hEb) <= tE * np.abs(hEb))
        assert np.all(np.abs(hRa - hRb) <= tR * max(np.abs(hRb).max(), 1e-300))
        assert H.rel(Ea, Eb) < tE
        assert dF <= tF and dV <= tF
        a.close()
        b.close()
    if name == "Chiral magnet":
        assert ran_tile


def test_fused41_survives_precision_switch_and_autotune():
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 128, 160)
    p_i, p_f = H.packed(th, p)
    F = H.chiral_field(th, p_i, p_f, seed=2)
    a = _sim(th, p, F, precision="fp32")
    info = a.autotune_tiles(reps=2)
    # 9 variants; FP32 chiral: the 32-row two-field tile does not fit LDS, the
    # 16- and 8-row tiles and every strip do
    assert len(info["ms_fused41"]) == 9 and info["ms_fused41"][0] == 0 and np.all(info["ms_fused41"][1:] > 0)
    assert len(info["ms_fused23"]) == 9 and info["ms_fused23"][0] == 0 and np.all(info["ms_fused23"][1:] > 0)
    for key in ("ms_pairs12", "ms_pairs34"):  # the paired sweeps run in the strip variants only
        assert len(info[key]) == 9 and np.all(info[key][:3] == 0) and np.all(info[key][3:] > 0)
    a.set_tile_variants([0, 3, 3, 3])
    a.set_fused_stages(True, True)
    b = _sim(th, p, F, precision="fp32")
    b.set_tile_variants([0, 3, 3, 3])
    E0 = a.compute_energy()
    Ea, _, _ = a.steps(E0, 5)
    Eb, _, _ = b.steps(E0, 5)
    a.set_precision("fp64")
    b.set_precision("fp64")
    assert a.tile_variants()["fused41"]
    Ea, _, _ = a.steps(Ea, 6)
    Eb, _, _ = b.steps(Eb, 6)
    # FP32 steps of the two schedules agree to FP32 rounding, which the FP64 steps carry on
    assert H.rel(Ea, Eb) < 1e-7
    assert np.abs(a.get_field() - b.get_field()).max() <= 1e-5
    a.close()
    b.close()


def test_fused_stages_rejected_when_two_rings_do_not_fit():
    th = ssa.load_theory("Spin triplet superconducting ferromagnet")
    p = H.st_params(th, 40, 40)
    p_i, p_f = H.packed(th, p)
    sim = _sim(th, p, H.st_field(th, p_i, p_f))
    with pytest.raises(ValueError, match="cannot share one sweep"):
        sim.set_fused_stages(True)
    with pytest.raises(ValueError, match="cannot share one sweep"):
        sim.set_fused_stages(False, True)
    sim.close()


_AUX_CASES = {
    "Ginzburg-Landau superconductor": (lambda th: H.gl_params(th, 50, 42), lambda th, pi, pf: H.gl_field(th, pi, pf, seed=2),
                                       {"higgs": (2, 3), "current": (0, 1, None)}),
    "Anyon superconductor": (lambda th: H.mcsh_params(th, 50, 42), lambda th, pi, pf: H.mcsh_field(th, pi, pf, seed=2),
                             {"higgs": (2, 3), "current": (0, 1, None)}),
    "Bose-Einstein condensate": (lambda th: H.bec_params(th, 50, 42), lambda th, pi, pf: H.bec_field(th, pi, pf, seed=2),
                                 {"higgs": (0, 1)}),
    "Anisotropic superconductor": (lambda th: H.as_params(th, 50, 42, seed=2),
                                   lambda th, pi, pf: H.as_field(th, pi, pf, seed=2, amp=5e-2),
                                   {"higgs1": (2, 3), "higgs2": (4, 5), "phase": True, "current": (0, 1, None),
                                    "field": [2]}),
    "Ferromagnetic superconductor": (lambda th: H.fs_params(th, 50, 42), lambda th, pi, pf: H.fs_field(th, pi, pf, seed=2),
                                     {"higgs": (6, 7), "current": (3, 4, 5), "field": [0, 1, 2]}),
    "Spin triplet superconducting ferromagnet": (lambda th: H.st_params(th, 50, 42),
                                                 lambda th, pi, pf: H.st_field(th, pi, pf, seed=2),
                                                 {"higgs1": (6, 7), "higgs2": (8, 9), "current": (3, 4, 5),
                                                  "field": [0, 1, 2]}),
}


@pytest.mark.parametrize("name", list(_AUX_CASES))
def test_aux_densities_match_oracle(name):
    """Higgs norms, phase difference, supercurrent and magnetic field (the
    reference Simulation's auxiliary densities) against the oracle."""
    from oracle import aux as oaux
    mkp, mkf, spec = _AUX_CASES[name]
    th = ssa.load_theory(name)
    p = mkp(th)
    p_i, p_f = H.packed(th, p)
    F = mkf(th, p_i, p_f)
    o = H.ORACLES[name]
    sim = _sim(th, p, F)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    tol = lambda ref: 1e-12 * max(1.0, float(np.abs(ref).max()))
    if "higgs" in spec:
        ref = oaux.higgs(F, *spec["higgs"])
        assert np.abs(sim.compute_higgs_density() - ref).max() <= tol(ref)
    for key, meth in (("higgs1", sim.compute_higgs1_density), ("higgs2", sim.compute_higgs2_density)):
        if key in spec:
            ref = oaux.higgs(F, *spec[key])
            assert np.abs(meth() - ref).max() <= tol(ref)
    if "phase" in spec:
        ref = oaux.phase_difference(F)
        assert np.abs(sim.compute_phase_difference_density() - ref).max() <= 1e-12
    if "current" in spec:
        ref = oaux.supercurrent(D2, *spec["current"])
        got = sim.supercurrent()
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-11 * max(1.0, float(np.abs(ref).max()))
    else:
        with pytest.raises(AttributeError):
            sim.supercurrent()
    if "field" in spec:
        ref = np.stack([o.vortex_density(F, D1, p_i, p_f, w) if o.vortex_density.__code__.co_argcount > 4
                        else o.vortex_density(F, D1, p_i, p_f) for w in spec["field"]])
        got = sim.magnetic_field()
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, float(np.abs(ref).max()))
    sim.close()


_BUNDLE_CASES = {
    "Chiral magnet": (lambda th: H.chiral_params(th, 30, 26), lambda th, pi, pf: H.chiral_field(th, pi, pf, seed=1)),
    "Ginzburg-Landau superconductor": (lambda th: H.gl_params(th, 30, 26), lambda th, pi, pf: H.gl_field(th, pi, pf)),
    "Baby Skyrme model": (lambda th: H.bs_params(th, 30, 26), lambda th, pi, pf: H.bs_field(th, pi, pf)),
    "Anisotropic superconductor": (lambda th: H.as_params(th, 30, 26), lambda th, pi, pf: H.as_field(th, pi, pf)),
    "Bose-Einstein condensate": (lambda th: H.bec_params(th, 30, 26), lambda th, pi, pf: H.bec_field(th, pi, pf)),
    "Anyon superconductor": (lambda th: H.mcsh_params(th, 30, 26), lambda th, pi, pf: H.mcsh_field(th, pi, pf)),
    "Ferromagnetic superconductor": (lambda th: H.fs_params(th, 30, 26), lambda th, pi, pf: H.fs_field(th, pi, pf)),
    "Spin triplet superconducting ferromagnet": (lambda th: H.st_params(th, 30, 26),
                                                 lambda th, pi, pf: H.st_field(th, pi, pf)),
    "Liquid crystal": (lambda th: H.lc_params(th, 30, 26, depol=True), lambda th, pi, pf: H.lc_field(th, pi, pf)),
}


@pytest.mark.parametrize("name", list(_BUNDLE_CASES))
def test_save_output_bundle_and_resume(tmp_path, name):
    """save_output writes the theory's file set (reference names); the energy
    density file holds the oracle's per-site energy; load_output restores the
    field exactly."""
    from oracle import io_format as oio
    mkp, mkf = _BUNDLE_CASES[name]
    th = ssa.load_theory(name)
    p = mkp(th)
    p_i, p_f = H.packed(th, p)
    F = mkf(th, p_i, p_f)
    sim = _sim(th, p, F)
    out = sim.save_output(str(tmp_path), precision=32)
    names = set(os.listdir(out))
    expected = {fn for _, _, fn in th.io.BUNDLE_SPEC} | {"Field.dat", "LatticePoints.dat", "LatticeVectors.dat"}
    assert names == expected
    o = H.ORACLES[name]
    e_ref = o.energy_density(F, ocore.first_derivatives(F, p_i, p_f), p_i, p_f)
    got = np.array(open(os.path.join(out, "EnergyDensity.dat")).read().split(), dtype=np.float64).reshape(e_ref.shape)
    assert np.abs(got - e_ref).max() <= 1e-12 * max(1.0, np.abs(e_ref).max())
    assert open(os.path.join(out, "Field.dat")).read() == oio.field_dat_text(F.ravel(), F.shape[1], F.shape[2],
                                                                              F.shape[0])
    sim2 = _sim(th, p, np.zeros_like(F))
    sim2.load_output(out)
    assert np.array_equal(sim2.get_field(), F)
    sim.close()
    sim2.close()


def test_observables_accept_the_reference_call_convention():
    """theory.observables functions called as the reference calls them
    (Field and scratch buffers first, e.g. chiral_magnet/observables.py:15-99,
    ferromagnetic_superconductor/observables.py compute_vortex_number(..., which, ...))."""
    th = ssa.load_theory("Ferromagnetic superconductor")
    p = H.fs_params(th, 40, 36)
    p_i, p_f = H.packed(th, p)
    sim = _sim(th, p, H.fs_field(th, p_i, p_f, seed=3))
    ob = th.observables
    E = ob.compute_energy(sim.Field, sim.d1fd1x, sim.en, sim.entmp, sim.gridsum_partial, sim.p_i_d, sim.p_f_d,
                          sim.p_i_h, sim.p_f_h)
    assert E == sim.compute_energy()
    Q = ob.compute_skyrmion_number(sim.Field, sim.d1fd1x, sim.en, sim.entmp, sim.gridsum_partial, sim.p_i_d,
                                   sim.p_f_d, sim.p_i_h)
    assert Q == ob.compute_skyrmion_number(sim)
    for which in (0, 1, 2):
        ref_style = ob.compute_vortex_number(sim.Field, sim.d1fd1x, sim.en, sim.entmp, sim.gridsum_partial, which,
                                             sim.p_i_d, sim.p_f_d, sim.p_i_h)
        assert ref_style == ob.compute_vortex_number(sim, which)
    with pytest.raises(TypeError):
        ob.compute_energy(np.zeros(3))
    sim.close()


@pytest.mark.parametrize("precision", ["fp32c", "fp64"])
def test_two_contexts_from_the_same_seed_agree_bit_for_bit(precision):
    """bench.py times its roofline on a second context that repeats the timed
    run's set-up (the same seeded field, launch shapes and warm-up) and then
    runs the timed steps eagerly, pass by pass.  That second run ends on the
    same energies, field and velocity as the first, bit for bit (paired
    sweeps; graphs of 4 steps in the first)."""
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 256, 240)
    cfg = {"mode": "initial", "ansatz": "anti", "noise": 1e-3, "seed": 5}

    def setup(graphs):
        s = ssa.Simulation(p, th, precision=precision)
        s.initialize(cfg)
        s.set_tile_variants([3, 3, 3, 3])
        s.set_fused_stages(False, False, pairs=True)
        if graphs:
            s.use_graphs(graphs)
        E = s.compute_energy()
        E, _, _ = s.steps(E, 6)  # warm-up
        return s, E

    a, Ea0 = setup(4)
    Ea, _, _ = a.steps(Ea0, 12, history=True)
    ha = a.last_history[0]
    b, Eb0 = setup(4)
    b.time_passes(Eb0, 12)
    hb = b.history(12)[0]
    rises = int(np.sum(np.diff(np.concatenate([[Ea0], ha])) > 0))
    print(f"[two contexts {precision}] energy rises={rises}")
    assert Ea0 == Eb0
    assert np.array_equal(ha, hb) and hb[-1] == Ea
    assert np.array_equal(a.get_field(), b.get_field())
    assert np.array_equal(a.get_velocity(), b.get_velocity())
    a.close()
    b.close()
