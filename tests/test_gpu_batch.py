"""Batched form (many lattices per launch) vs single-lattice Simulations and the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import soliton_solver_amd as ssa
from oracle import core as ocore
from soliton_solver_amd.batch import BatchSimulation
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _bs_instances(sizes, seed=0):
    th = ssa.load_theory("Baby Skyrme model")
    rng = np.random.default_rng(seed)
    params, cfgs = [], []
    for k, n in enumerate(sizes):
        mpi = rng.uniform(0.1, 1.0)
        kappa = rng.uniform(0.5, 2.0)
        params.append(H.bs_params(th, n, n + 8 * (k % 3), mpi=mpi, kappa=kappa))
        cfgs.append({"mode": "initial", "ansatz": "neel", "skyrmion_number": int(rng.integers(1, 5)),
                     "noise": 1e-3, "seed": seed * 100 + k})
    return th, params, cfgs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_matches_individual_simulations(precision):
    th, params, cfgs = _bs_instances([32, 48, 64, 40, 72], seed=1)
    batch = BatchSimulation(th, params, precision=precision)
    batch.initialize(cfgs)
    n_active = batch.run(30, tol=0.0, check_every=10)
    assert n_active == len(params)
    st = batch.state()
    for i, (p, cfg) in enumerate(zip(params, cfgs)):
        sim = ssa.Simulation(p, th, precision=precision)
        sim.initialize(cfg)
        E, _, ea = sim.steps(sim.compute_energy(), 30)
        Fb = batch.get_field(i)
        Fs = sim.get_field()
        tolE = 1e-12 if precision == "fp64" else 1e-6
        tolF = 1e-10 if precision == "fp64" else 1e-5
        print(f"[batch {precision} inst {i} {p.xlen}x{p.ylen}] E_b={st['energy'][i]:.15e} E_s={E:.15e} "
              f"max|dF|={np.abs(Fb - Fs).max():.2e}")
        assert abs(st["energy"][i] - E) <= tolE * abs(E)
        assert np.abs(Fb - Fs).max() <= tolF
        assert st["steps"][i] == 30
        sim.close()
    batch.close()


def test_batch_instance_matches_oracle():
    th, params, cfgs = _bs_instances([36, 44], seed=2)
    batch = BatchSimulation(th, params, precision="fp64")
    batch.initialize(cfgs)
    F0 = [batch.get_field(i).copy() for i in range(2)]
    batch.run(15, tol=0.0)
    st = batch.state()
    o = H.ORACLES["Baby Skyrme model"]
    for i in range(2):
        p_i, p_f = H.packed(th, params[i])
        F = F0[i].copy()
        V = np.zeros_like(F)
        Es, _ = ocore.relax(o, F, V, p_i, p_f, 15)
        assert abs(st["energy"][i] - Es[-1]) <= 1e-10 * abs(Es[-1])
        assert np.abs(batch.get_field(i) - F).max() <= 1e-9
    batch.close()


def test_batch_graphs_replay_like_eager_and_instances_retire():
    th, params, cfgs = _bs_instances([32, 64, 48, 80], seed=3)
    a = BatchSimulation(th, params, precision="fp32")
    b = BatchSimulation(th, params, precision="fp32")
    a.initialize(cfgs)
    b.initialize(cfgs)
    b.use_graphs(8)
    tol = 5e-3
    for _ in range(5):
        a.run(16, tol=tol, check_every=16)
        b.run(16, tol=tol, check_every=16)
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa["energy"], sb["energy"])
    assert np.array_equal(sa["steps"], sb["steps"])
    for i in range(len(params)):
        assert np.array_equal(a.get_field(i), b.get_field(i))
    # a retired instance does not move any more
    conv = sa["converged_step"]
    if (conv >= 0).any():
        i = int(np.argmax(conv >= 0))
        F_before = a.get_field(i).copy()
        a.run(10, tol=tol)
        assert np.array_equal(a.get_field(i), F_before)
        assert a.state()["steps"][i] == sa["steps"][i]
    a.close()
    b.close()
