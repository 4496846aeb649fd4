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


_THEORY_BATCHES = {
    "Chiral magnet": ((lambda th, n, m: H.chiral_params(th, n, m)),
                      (lambda th, pi, pf, k: H.chiral_field(th, pi, pf, seed=10 + k, rough=True))),
    "Ginzburg-Landau superconductor": ((lambda th, n, m: H.gl_params(th, n, m)),
                                       (lambda th, pi, pf, k: H.gl_field(th, pi, pf, seed=10 + k))),
    # xsize 16: at the default 25 the coarser 40-row lattice (h = 0.64) blows up within 8 steps, oracle included
    "Anisotropic superconductor": ((lambda th, n, m: H.as_params(th, n, m, seed=3, xsize=16.0)),
                                   (lambda th, pi, pf, k: H.as_field(th, pi, pf, seed=10 + k))),
    "Anyon superconductor": ((lambda th, n, m: H.mcsh_params(th, n, m)),
                             (lambda th, pi, pf, k: H.mcsh_field(th, pi, pf, seed=10 + k))),
    "Ferromagnetic superconductor": ((lambda th, n, m: H.fs_params(th, n, m, eta2=0.05)),
                                     (lambda th, pi, pf, k: H.fs_field(th, pi, pf, seed=10 + k))),
    "Spin triplet superconducting ferromagnet": ((lambda th, n, m: H.st_params(th, n, m)),
                                                 (lambda th, pi, pf, k: H.st_field(th, pi, pf, seed=10 + k))),
    "Liquid crystal": ((lambda th, n, m: H.lc_params(th, n, m)),
                       (lambda th, pi, pf, k: H.lc_field(th, pi, pf, seed=10 + k))),
}


@pytest.mark.parametrize("name", list(_THEORY_BATCHES))
def test_batch_of_each_theory_matches_oracle(name):
    """Two instances of different shapes of every batchable theory in one
    batch (FP64): each follows the oracle's trajectory (energy 1e-10
    relative, max |dF| 1e-9) over 8 steps."""
    th = ssa.load_theory(name)
    mkp, mkf = _THEORY_BATCHES[name]
    params = [mkp(th, 40, 52), mkp(th, 57, 36)]
    fields = []
    for k, p in enumerate(params):
        p_i, p_f = H.packed(th, p)
        fields.append(mkf(th, p_i, p_f, k))
    batch = BatchSimulation(th, params, precision="fp64")
    batch.initialize([{"mode": "array", "field": F} for F in fields])
    batch.run(8, tol=0.0)
    st = batch.state()
    o = H.ORACLES[name]
    for i, (p, F0) in enumerate(zip(params, fields)):
        p_i, p_f = H.packed(th, p)
        F = F0.copy()
        V = np.zeros_like(F)
        Es, _ = ocore.relax(o, F, V, p_i, p_f, 8)
        dF = np.abs(batch.get_field(i) - F).max()
        print(f"[batch {name} inst {i}] relE={abs(st['energy'][i] - Es[-1]) / abs(Es[-1]):.2e} max|dF|={dF:.2e}")
        assert abs(st["energy"][i] - Es[-1]) <= 1e-10 * abs(Es[-1])
        assert dF <= 1e-9
    batch.close()


def test_batch_autotune_keeps_state_and_trajectory():
    """Autotuning the batch's tile shape leaves the state alone, and a run in
    the chosen shape follows the default-shape run (FP64: 1e-12 relative
    energy, fields to 1e-12; only partial-sum order differs)."""
    th, params, cfgs = _bs_instances([48, 100, 64, 130], seed=7)
    a = BatchSimulation(th, params, precision="fp64")
    b = BatchSimulation(th, params, precision="fp64")
    a.initialize(cfgs)
    b.initialize(cfgs)
    a.run(5, tol=0.0)
    b.run(5, tol=0.0)
    F0 = [a.get_field(i).copy() for i in range(len(params))]
    info = a.autotune_tiles(reps=3)
    assert 0 <= info["variant"] < 9 and len(info["ms"]) == 9 and np.all(info["ms"] > 0)
    assert len(info["ms_fused23"]) == 9 and np.all(info["ms_fused23"][3:] > 0)  # every strip has the 2+3 sweep
    assert len(info["ms_pairs"]) == 9 and np.all(info["ms_pairs"][:3] == 0) and np.all(info["ms_pairs"][3:] > 0)
    for i in range(len(params)):
        assert np.array_equal(a.get_field(i), F0[i])
    a.run(4, tol=0.0)
    b.run(4, tol=0.0)
    for v in (2, 1, 0, 3, 4, 5, 6, 7, 8):
        a.set_tile_variant(v)
        assert a.tile_variant()["variant"] == v
        a.run(4, tol=0.0)
        b.run(4, tol=0.0)
    sa, sb = a.state(), b.state()
    for i in range(len(params)):
        assert abs(sa["energy"][i] - sb["energy"][i]) <= 1e-12 * abs(sb["energy"][i])
        assert np.abs(a.get_field(i) - b.get_field(i)).max() <= 1e-12
        assert sa["steps"][i] == sb["steps"][i] == 45
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_pairs_match_separate_passes(precision):
    """The paired sweeps 1+2 and 3+4 over every instance (strips; passes 1+2
    read V before each instance's arrest decision, passes 3+4 take l2 = l1
    where it arrests) against the separate passes: the same energies, fields
    and step counts after 12 steps; tile variants have no paired sweeps."""
    th, params, cfgs = _bs_instances([48, 100, 64, 130], seed=11)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for v in (3, 6):
        a = BatchSimulation(th, params, precision=precision)
        b = BatchSimulation(th, params, precision=precision)
        a.initialize(cfgs)
        b.initialize(cfgs)
        a.set_tile_variant(v)
        b.set_tile_variant(v)
        a.set_fused_stages(False, pairs=True)
        assert a.tile_variant()["pairs"] and not b.tile_variant()["pairs"]
        a.run(12, tol=0.0)
        b.run(12, tol=0.0)
        sa, sb = a.state(), b.state()
        for i in range(len(params)):
            dF = np.abs(a.get_field(i) - b.get_field(i)).max()
            print(f"[batch pairs {precision} v={v} inst {i}] relE="
                  f"{abs(sa['energy'][i] - sb['energy'][i]) / abs(sb['energy'][i]):.2e} max|dF|={dF:.2e}")
            assert abs(sa["energy"][i] - sb["energy"][i]) <= tol * abs(sb["energy"][i])
            assert dF <= tol
            assert sa["steps"][i] == sb["steps"][i] == 12
        a.close()
        b.close()
    with pytest.raises(ValueError):
        c = BatchSimulation(th, params, precision=precision)
        try:
            c.set_tile_variant(1)
            c.set_fused_stages(False, pairs=True)  # tiles: no paired sweeps
        finally:
            c.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_fused23_matches_separate_passes(precision):
    """Passes 2+3 of every instance in one sweep (tiles where the two-field tile
    fits, strips) against the separate passes: the same energies and fields."""
    th, params, cfgs = _bs_instances([48, 100, 64, 130], seed=11)
    tol = 1e-12 if precision == "fp64" else 1e-5
    ran = []
    for v in (1, 2, 3, 6):
        a = BatchSimulation(th, params, precision=precision)
        b = BatchSimulation(th, params, precision=precision)
        a.initialize(cfgs)
        b.initialize(cfgs)
        a.set_tile_variant(v)
        b.set_tile_variant(v)
        try:
            a.set_fused_stages(True)
        except ValueError:  # no two-field tile of this variant fits
            a.close()
            b.close()
            continue
        assert a.tile_variant()["fused23"] and not b.tile_variant()["fused23"]
        ran.append(v)
        a.run(12, tol=0.0)
        b.run(12, tol=0.0)
        sa, sb = a.state(), b.state()
        for i in range(len(params)):
            dF = np.abs(a.get_field(i) - b.get_field(i)).max()
            assert abs(sa["energy"][i] - sb["energy"][i]) <= tol * abs(sb["energy"][i])
            assert dF <= tol
            assert sa["steps"][i] == sb["steps"][i] == 12
        a.close()
        b.close()
    print(f"[batch fused23 {precision}] variants run fused: {ran}")
    assert 3 in ran and 6 in ran


def test_batch_precision_switch_keeps_the_state():
    """ss_batch_set_precision: FP32 -> FP64 converts every instance's field and
    velocity exactly (FP32 values are FP64 values), and FP64 -> FP32 rounds them."""
    th, params, cfgs = _bs_instances([40, 56, 48], seed=4)
    batch = BatchSimulation(th, params, precision="fp32")
    try:
        batch.initialize(cfgs)
        batch.run(20, tol=0.0)
        F32 = [batch.get_field(i) for i in range(3)]
        batch.set_precision("fp64")
        assert batch.precision == "fp64"
        for i in range(3):
            assert np.array_equal(batch.get_field(i), F32[i])
        batch.run(10, tol=0.0)  # steps run in FP64 after the switch
        F64 = [batch.get_field(i) for i in range(3)]
        batch.set_precision("fp32")
        for i in range(3):
            assert np.abs(batch.get_field(i) - F64[i]).max() <= 1e-7 * max(1.0, np.abs(F64[i]).max())
    finally:
        batch.close()


def test_batch_fp64_retire_phase_reaches_1e_6():
    """BatchSimulation.relax: FP32 to the FP32 floor (1e-3), then the FP64
    retire phase to max |EG| < 1e-6, against single-lattice FP64 relaxations
    of the same charge-1 fields to the same tolerance: same energy to 1e-8
    relative (both are at the minimum; the FP32 phase only changes the path)."""
    th = ssa.load_theory("Baby Skyrme model")
    params = [H.bs_params(th, n, n, mpi=0.8, kappa=1.0) for n in (64, 80, 72)]
    cfgs = [{"mode": "initial", "ansatz": "neel", "skyrmion_number": 1, "noise": 1e-3, "seed": 30 + k}
            for k in range(3)]
    batch = BatchSimulation(th, params, precision="fp32")
    try:
        batch.initialize(cfgs)
        res = batch.relax(1e-6, 120000, check_every=500, reduced_steps=20000)
        print(f"[batch retire] converged {res['converged']}/3 at steps {res['converged_step'].tolist()} "
              f"(fp32 {res['steps_reduced'].tolist()}, fp64 {res['steps_fp64'].tolist()}) "
              f"err {res['err_absmax'].tolist()}", flush=True)
        assert res["converged"] == 3 and np.all(res["err_absmax"] < 1e-6)
        for i, (p, cfg) in enumerate(zip(params, cfgs)):
            sim = ssa.Simulation(p, th)
            sim.initialize(cfg)
            E, ea, n = sim.minimize(tol=1e-6, max_steps=200000, check_every=500)
            print(f"[batch retire inst {i}] batch E={res['energy'][i]:.13f} single FP64 E={E:.13f} ({n} steps) "
                  f"relE={H.rel(res['energy'][i], E):.2e}", flush=True)
            assert ea < 1e-6
            assert H.rel(res["energy"][i], E) < 1e-8
            sim.close()
    finally:
        batch.close()
