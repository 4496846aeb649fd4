"""Failure detection: a non-finite field must surface in the energy and the
error of a step and must never read as converged (the reference's
reduce_max_kernel / numpy max keep a NaN, core/utils.py:230-271, and its loop
compares err with the tolerance, core/integrator.py:338-347).  The NaN is put
into the field from the host (an ordinary upload); no kernel faults."""
from __future__ import annotations

import math

import numpy as np
import pytest

import soliton_solver_amd as ssa
from tests.helpers import chiral_field, chiral_params, packed

pytestmark = pytest.mark.gpu


def _sim(precision, n=48, pairs=False):
    th = ssa.load_theory("Chiral magnet")
    p = chiral_params(th, n=n)
    sim = ssa.Simulation(p, th, precision=precision)
    p_i, p_f = packed(th, p)
    F = chiral_field(th, p_i, p_f, seed=3)
    return th, sim, F


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
@pytest.mark.parametrize("schedule", ["separate", "paired"])
def test_nan_site_surfaces_in_energy_and_error(precision, schedule):
    th, sim, F = _sim(precision, n=96 if schedule == "paired" else 48)
    try:
        if schedule == "paired":
            sim.set_tile_variants([3, 3, 3, 3])
            sim.set_fused_stages(False, False, pairs=True)
        F[1, 30, 17] = np.nan
        sim.set_field(F)
        E, err, ea = sim.steps(1.0e300, 3)
        print(f"[nan] {precision} {schedule}: E={E} err={err} err_abs={ea}", flush=True)
        assert math.isnan(ea)
        assert not math.isfinite(E)
        with pytest.raises(FloatingPointError):
            sim.minimize(tol=1e300, max_steps=4, check_every=2)
    finally:
        sim.close()


def test_infinite_site_never_converges():
    th, sim, F = _sim("fp64")
    try:
        F[0, 20, 20] = np.inf
        sim.set_field(F)
        E, err, ea = sim.steps(1.0e300, 2)
        assert not (ea < 1e300)
        with pytest.raises(FloatingPointError):
            sim.minimize(tol=1e300, max_steps=2, check_every=1)
    finally:
        sim.close()


def test_batch_instance_with_nan_retires_as_failed():
    from soliton_solver_amd.batch import BatchSimulation
    th = ssa.load_theory("Baby Skyrme model")
    params = [th.params.default_params(xlen=n, ylen=n, xsize=20.0, ysize=20.0, mpi=0.5, kappa=1.0,
                                       unit_magnetization=True, time_step=0.0957 * 20.0 / (n - 1))
              for n in (64, 96, 64)]
    bs = BatchSimulation(th, params, precision="fp32")
    try:
        cfgs = [{"mode": "initial", "ansatz": "neel", "skyrmion_number": 1, "noise": 1e-3, "seed": i}
                for i in range(3)]
        bs.initialize(cfgs)
        F = bs.get_field(1)
        F[2, 40, 41] = np.nan
        bs.set_field(1, F)
        bs.begin()
        active = bs.run(40, tol=1e-30, check_every=20)
        st = bs.state()
        print(f"[nan] batch: active={active} conv={st['converged_step']} err={st['err_absmax']}", flush=True)
        assert st["failed"].tolist() == [False, True, False]
        assert st["converged_step"][1] == -2
        assert math.isnan(st["err_absmax"][1])
        assert active == 2
        assert np.all(np.isfinite(st["err_absmax"][[0, 2]]))
    finally:
        bs.close()


def test_device_pointer_calls_need_fp64():
    torch = pytest.importorskip("torch")
    from soliton_solver_amd import _lib
    th, sim, F = _sim("fp32")
    try:
        sim.set_field(F)
        x = torch.zeros(F.size, dtype=torch.float64, device="cuda")
        g = torch.zeros_like(x)
        rc = _lib.lib().ss_variation_dev(sim._ctx, x.data_ptr(), g.data_ptr(), 0, None)
        assert rc == -1 and "FP64" in _lib.error_text(sim._ctx)
        e = torch.zeros(1, dtype=torch.float64, device="cuda")
        rc = _lib.lib().ss_energy_dev(sim._ctx, x.data_ptr(), e.data_ptr(), None)
        assert rc == -1
    finally:
        sim.close()
