"""End-to-end relaxations of the two single-lattice FP64 workloads (GPU).

* Abelian-Higgs vortex, 128 x 128, spacing 0.2, lambda = 1, u = 1 (critical
  coupling, E -> pi), psi = tanh(r) e^{i theta} + U(-1e-3, 1e-3), A = 0,
  2000 arrested-Newton steps at dt = 0.1: the GPU trajectory against the
  oracle's (energy 1e-10 relative, max |dF| 1e-8), and E, N against pi, -1.
* Chiral magnet, 320 x 320, xsize 10, Heusler DMI, demag, one anti-skyrmion
  of radius 1 at the centre with N(0, 1e-6) noise: relaxed to max |EG| < 1e-8,
  skyrmion number +-1.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import soliton_solver_amd as ssa
from oracle import core as ocore
from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_abelian_higgs_vortex_128_relaxes_to_pi():
    th = ssa.load_theory("Abelian Higgs")
    p = H.gl_params(th, 128, 128, lsx=0.2, time_step=0.1)
    p_i, p_f = H.packed(th, p)
    F = np.zeros((4, 128, 128))
    th.initial_config.initialize(Field=F, Velocity=np.zeros_like(F), grid=H.grid_of(p_i, p_f), p_i_h=p_i,
                                 p_f_h=p_f, config={"mode": "vortex", "noise": 1e-3, "seed": 0})
    o = H.ORACLES["Ginzburg-Landau superconductor"]
    Fo, Vo = F.copy(), np.zeros_like(F)
    E0o = ocore.theory_energy(o, Fo, p_i, p_f)
    Es, _ = ocore.relax(o, Fo, Vo, p_i, p_f, 2000, prev_energy=E0o)
    sim = ssa.Simulation(p, th)
    sim.initialize({"mode": "array", "field": F})
    E0 = sim.compute_energy()
    E, err, ea = sim.steps(E0, 2000)
    N = sim.observables()["vortex_number"]
    dF = np.abs(sim.get_field() - Fo).max()
    print(f"[abelian higgs 128] E={E:.12f} (pi={math.pi:.12f}) N={N:.10f} max|EG|={ea:.2e} "
          f"relE vs oracle={H.rel(E, Es[-1]):.2e} max|dF|={dF:.2e}")
    assert H.rel(E, Es[-1]) < 1e-10
    assert dF < 1e-8
    assert abs(E - math.pi) / math.pi < 5e-3
    # the coupling q A.(psi1 grad psi2 - psi2 grad psi1) makes the flux of a
    # psi ~ e^{i theta} vortex negative: N = -1
    assert abs(N + 1.0) < 1e-3
    sim.close()


def test_chiral_320_relaxes_below_1e_8():
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, 320, 320, xsize=10.0)
    p_i, p_f = H.packed(th, p)
    sim = ssa.Simulation(p, th)
    sim.initialize({"mode": "skyrmions", "ansatz": "anti", "skyrmions": [(5.0, 5.0, 1.0)], "noise": 1e-6,
                    "seed": 1})
    E0 = sim.compute_energy()
    sim.use_graphs(20)
    # the demagnetising potential is a saddle field: the energy rises while it
    # relaxes, so the flow's velocity is killed nearly every step and the last
    # decades take ~2e5 steps (measured: 6.9e-9 after 2.1e5)
    E, ea, n = sim.minimize(tol=1e-8, max_steps=400000, check_every=5000)
    Q = sim.observables()["skyrmion_number"]
    print(f"[chiral 320] E0={E0:.10f} E={E:.12f} Q={Q:.10f} steps={n} max|EG|={ea:.2e}")
    assert ea < 1e-8
    assert abs(abs(Q) - 1.0) < 1e-2
    sim.close()
