"""CPU checks of the boundary: the C-ABI library loads and exports every symbol
include/soliton.h declares; host-side parameter packing and the registry
behave like the reference's (theories/registry.py, core/params.py)."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import soliton_solver_amd as ssa
from soliton_solver_amd import _lib
from soliton_solver_amd.theories import registry

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "soliton.h").read_text()
    return sorted(set(re.findall(r"\b(ss_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ss_\w+)", out))
    assert set(names) <= exported
    assert set(_lib.SIGNATURES) == set(names)


def test_library_reports_version_and_no_crash_without_gpu():
    L = _lib.lib()
    assert L.ss_version() >= 100
    assert _lib.device_count() >= 0


def test_registry_lookup_and_aliases():
    assert "Chiral magnet" in ssa.list_theories()
    a = ssa.load_theory("magnetic skyrmion")
    b = ssa.load_theory("  CHIRAL MAGNET ")
    assert a is b
    assert ssa.load_theory("Abelian Higgs").THEORY_SPEC.name == "Ginzburg-Landau superconductor"
    with pytest.raises(KeyError):
        ssa.load_theory("no such theory")
    with pytest.raises(ValueError):
        registry.register_theory(registry.TheorySpec(name="Chiral magnet", import_path="x"))
    assert registry.DISCOVERY_ERRORS == {}


def test_core_param_resolution():
    from soliton_solver_amd.core.params import default_params, pack_device_params
    rp = default_params(xlen=11, ylen=21, xsize=5.0, ysize=4.0, number_total_fields=3).resolved()
    assert rp.lsx == 0.5 and rp.lsy == 0.2 and rp.grid_volume == 0.5 * 0.2
    assert rp.time_step == 0.25  # courant * lsx, lsy plays no part
    p_i, p_f = pack_device_params(rp)
    assert p_i.dtype == np.int32 and p_f.dtype == np.float64
    assert list(p_i) == [11, 21, 2, 2, 3, 231, 693, 1, 1, 0]
    assert list(p_f) == [5.0, 4.0, 0.5, 0.2, 0.1, 0.25]
    rp2 = default_params(lsx=0.3, time_step=0.01).resolved()
    assert rp2.lsx == 0.3 and rp2.time_step == 0.01


def test_chiral_params_packing_and_dmi_selector():
    th = ssa.load_theory("Chiral magnet")
    p = th.params.default_params(xlen=320, ylen=320, xsize=10.0, ysize=10.0, dmi_term="Heusler", demag=True,
                                 ansatz="anti", unit_magnetization=True)
    p_i, p_f = th.params.pack_device_params(p.resolved())
    assert len(p_i) == 16 and len(p_f) == 15
    assert list(p_i[10:]) == [3, 0, 0, 1, 0, 1]
    assert abs(p_f[6] - 2.0) < 1e-12                      # coup_K = K J / D^2
    assert abs(p_f[8] - 1.25663706127e-6 * 580e3 ** 2 * 40e-12 / 16e-6) < 1e-12
    assert list(p_f[11:]) == [0.0, 0.0, 1.0, 0.0]         # anti ansatz flag
    assert th.params.default_params(dmi_term=["bulk", "b20"]).resolved().dmi_dresselhaus
    with pytest.raises(ValueError):
        th.params.default_params(dmi_term="weird").resolved()
    rp = th.params.default_params(D=0.0).resolved()
    assert (rp.coup_K, rp.coup_h, rp.coup_mu) == (1.0, 1.0, 1.0)


def test_gl_params_packing():
    th = ssa.load_theory("Ginzburg-Landau superconductor")
    p_i, p_f = th.params.pack_device_params(th.params.default_params(q=2.0, vortex_number=3.0).resolved())
    assert list(p_i[10:]) == [2, 2]
    assert list(p_f[6:]) == [2.0, 1.0, 1.0, 3.0, 1.5]     # ainf = N / q


def test_simulation_requires_gradient_kernel():
    import types
    th = ssa.load_theory("Chiral magnet")
    fake = types.SimpleNamespace(kernels=types.SimpleNamespace(), initial_config=th.initial_config,
                                 observables=th.observables, params=th.params)
    with pytest.raises(AttributeError):
        ssa.Simulation(th.params.default_params(xlen=16, ylen=16), fake)


def test_batch_refuses_instances_beyond_32bit_offsets():
    """The batched strip kernels address an instance's arrays (nf planes each)
    through buffer resources with 32-bit byte offsets below 2^31, in FP64 (the
    widest precision a batch converts to): nf * sites * 8 < 2^31.  A larger
    instance is refused when the batch is created, before any device is
    touched; 10000^2 baby-Skyrme sites (nf = 3: 2.4e9 bytes per array) are
    beyond it."""
    from soliton_solver_amd.batch import BatchSimulation
    th = ssa.load_theory("Baby Skyrme model")
    small = th.params.default_params(xlen=64, ylen=64, xsize=20.0, ysize=20.0)
    for n in (10000, 24000):
        huge = th.params.default_params(xlen=n, ylen=n, xsize=20.0, ysize=20.0)
        with pytest.raises(ValueError, match="exceed"):
            BatchSimulation(th, [small, huge], precision="fp32")


def test_bench_arrested_steps_follow_the_energy_history():
    """bench.py prices the dominant launch of each timed step at its
    algorithmic bytes, which drop when the step's velocity was arrested.  The
    arrest of step s >= 1 is E(F_s) > E(F_{s-1}) (core/integrator.py:338-343),
    with E(F_0) the energy the call starts from and history[k] = E(F_{k+1});
    step 0 starts from the caller's decision and counts as free."""
    from bench import arrested_steps
    hist = [5.0, 4.0, 4.5, 4.4, 4.6, 4.7]
    kills = arrested_steps(hist, 6.0)
    # E(F_1..F_5) = 5.0, 4.0, 4.5, 4.4, 4.6 against 6.0, 5.0, 4.0, 4.5, 4.4
    assert kills.tolist() == [False, False, False, True, False, True]
    assert arrested_steps([7.0], 6.0).tolist() == [False]


def test_bench_schedule_words_and_launch_pricing():
    """The words per site and component bench.py prices each schedule at.  The
    unfused 4-pass schedule moves 2 + 3 + 4 + 7 = 16 words per update, the
    fixed algorithmic count (16 nf w bytes); the paired sweeps move 4 + 6 (F, V
    -> l1, l2; F, V, l1, l2 -> F', V'), the compensated ones 5 + 8 (the low
    words: read in the 1+2 sweep, read and written in the 3+4 sweep).  An
    arrested 3+4 sweep reads no V and takes l2 = l1: 2 words fewer."""
    from bench import launch_words, schedule
    base = {"rows": [64, 64, 64, 64], "fused23": 0, "fused41": 0, "pairs": 0, "variants": [3, 3, 3, 3]}
    plain = schedule(base, False)
    assert (plain["step_words"], plain["dominant_words"], plain["arrested_words"]) == (16, 7, 6)
    pairs = schedule(dict(base, pairs=1), False)
    assert (pairs["step_words"], pairs["dominant_words"], pairs["arrested_words"]) == (10, 6, 4)
    assert pairs["kernel"] == "stream_pass<ChiralMagnet, float, 34, 1>"
    comp = schedule(dict(base, pairs=1, comp12=1, variants=[3, 6, 6, 6]), True)
    assert (comp["step_words"], comp["dominant_words"], comp["arrested_words"]) == (13, 8, 6)
    assert comp["kernel"] == "stream_pass_comp<ChiralMagnet, float, 34, 2>"
    # mean words over the launches: free steps at 8, arrested ones at 6
    assert launch_words(comp, [False, True, False, False]) == (3 * 8 + 6) / 4
    assert launch_words(comp, [False] * 5) == 8.0
    assert launch_words(comp, [True] * 3) == 6.0


def test_packed_parameter_jacobian_copies_are_exactly_one():
    """dp_f/dtheta of the anisotropic packing (host-side, no GPU): entries
    that copy a parameter are exactly 1 (the difference quotient divides by
    the step actually taken), the geometry rows are 0, and a derived entry
    matches a central difference at a tenth of the step."""
    import numpy as np
    import soliton_solver_amd as ssa
    from soliton_solver_amd import autograd as sag
    from tests import helpers as H
    th = ssa.load_theory("Anisotropic superconductor")
    p = H.as_params(th, 32, 32, xsize=10.0, seed=1)

    class _Sim:  # what packed_parameter_jacobian reads of a Simulation
        pass
    sim = _Sim()
    sim.p, sim.theory = p, th
    sim.p_f_h = np.asarray(th.params.pack_device_params(p.resolved())[1], dtype=np.float64)
    names = ["alpha1", "alpha2", "beta1", "beta2", "beta3", "beta4", "kappa"]
    J = sag.packed_parameter_jacobian(sim, names)
    assert np.all(J[:6] == 0.0)
    copies = np.abs(J - 1.0) < 1e-6
    assert copies.sum() >= len(names)  # every named parameter is packed at least once as itself
    assert np.all(J[copies] == 1.0)
    J10 = sag.packed_parameter_jacobian(sim, names, rel_step=1e-7)
    derived = (np.abs(J) > 1e-3) & ~copies
    assert derived.any()
    assert np.allclose(J[derived], J10[derived], rtol=1e-6)


def test_bench_theory_cases_initialise_and_step_in_the_oracle():
    """Each theory of the theories workload builds its parameters and initial
    condition and takes one finite step of the oracle's flow (32^2 lattice)."""
    import importlib
    import numpy as np
    import soliton_solver_amd as ssa
    from bench import THEORY_CASES, theory_params
    from oracle import core as ocore
    assert len(THEORY_CASES) == 9
    for name, oname, kw, xsize, dt_h, cfg in THEORY_CASES:
        th = ssa.load_theory(name)
        p_i, p_f = th.params.pack_device_params(theory_params(th, name, kw, xsize, dt_h, 32).resolved())
        p_i = np.ascontiguousarray(p_i, dtype=np.int32)
        p_f = np.ascontiguousarray(p_f, dtype=np.float64)
        X, Y, nf = int(p_i[0]), int(p_i[1]), int(p_i[4])
        g = np.empty((2, X, Y))
        g[0] = p_f[2] * np.arange(X)[:, None]
        g[1] = p_f[3] * np.arange(Y)[None, :]
        F = np.zeros((nf, X, Y))
        V = np.zeros_like(F)
        th.initial_config.initialize(Field=F, Velocity=V, grid=g, p_i_h=p_i, p_f_h=p_f, config=dict(cfg, seed=1))
        o = importlib.import_module(f"oracle.theories.{oname}")
        E0 = ocore.theory_energy(o, F, p_i, p_f)
        ocore.relax(o, F, V, p_i, p_f, 1, prev_energy=E0)
        assert np.isfinite(E0) and np.isfinite(F).all(), name
