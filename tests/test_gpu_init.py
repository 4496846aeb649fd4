"""Initial conditions generated on the device (ss_init_field, ss_add_noise)
against the host generators of each theory's ``initial_config`` on the same
configuration.

Noise-free fields agree with the host ones to rounding: max |F_dev - F_host|
<= 1e-11 * max(1, max|F_host|) in FP64 (transcendentals differ by an ulp or so
between the device math library and NumPy's), and within 1e-6 relative in
FP32 storage.  Velocity is zeroed.  Noise comes from a counter-based generator
on the device, so it is checked by its properties: deterministic per seed,
bounded (uniform) or of the stated spread (Gaussian), on the stated components
only, and the unit constraint restored where the host generator restores it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import soliton_solver_amd as ssa
from soliton_solver_amd import _lib
from tests import helpers as H

pytestmark = pytest.mark.gpu

_PARAMS = {
    "Chiral magnet": lambda th: H.chiral_params(th, 48, 40),
    "Baby Skyrme model": lambda th: H.bs_params(th, 52, 44),
    "Liquid crystal": lambda th: H.lc_params(th, 50, 42),
    "Ginzburg-Landau superconductor": lambda th: H.gl_params(th, 48, 40),
    "Anyon superconductor": lambda th: H.mcsh_params(th, 46, 40),
    "Anisotropic superconductor": lambda th: H.as_params(th, 48, 42),
    "Ferromagnetic superconductor": lambda th: H.fs_params(th, 50, 44),
    "Spin triplet superconducting ferromagnet": lambda th: H.st_params(th, 50, 44),
    "Bose-Einstein condensate": lambda th: H.bec_params(th, 48, 44),
}

_CASES = [
    ("Chiral magnet", {"mode": "ground"}),
    ("Chiral magnet", {"mode": "initial", "ansatz": "bloch"}),
    ("Chiral magnet", {"mode": "initial", "ansatz": "neel", "skyrmion_rotation": 0.4}),
    ("Chiral magnet", {"mode": "initial", "ansatz": "anti"}),
    ("Chiral magnet", {"mode": "skyrmions", "ansatz": "neel",
                       "skyrmions": [(2.6, 2.3, 0.9), (7.1, 5.9, 1.1)]}),
    ("Chiral magnet", {"mode": "skyrmions", "skyrmions": [(4.2, 4.0, 1.3), (5.9, 4.4, 1.3)]}),
    ("Baby Skyrme model", {"mode": "ground"}),
    ("Baby Skyrme model", {"mode": "initial", "ansatz": "neel", "skyrmion_number": 2}),
    ("Baby Skyrme model", {"mode": "skyrmions", "skyrmion_number": 0, "skyrmion_rotation": 0.2,
                           "skyrmions": [(6.0, 7.5, 2.0), (13.3, 9.1, 2.5)]}),
    ("Liquid crystal", {"mode": "ground"}),
    ("Liquid crystal", {"mode": "initial", "ansatz": "bloch"}),
    ("Liquid crystal", {"mode": "initial", "ansatz": "anti", "skyrmion_rotation": 1.0}),
    ("Ginzburg-Landau superconductor", {"mode": "ground"}),
    ("Ginzburg-Landau superconductor", {"mode": "initial", "vortex_number": 2.0}),
    ("Ginzburg-Landau superconductor", {"mode": "vortex", "vortex_number": -1.0, "scale": 0.7}),
    ("Anyon superconductor", {"mode": "ground"}),
    ("Anyon superconductor", {"mode": "initial", "vortex_number": 1.0}),
    ("Anisotropic superconductor", {"mode": "ground"}),
    ("Anisotropic superconductor", {"mode": "initial"}),
    ("Ferromagnetic superconductor", {"mode": "ground"}),
    ("Ferromagnetic superconductor", {"mode": "initial", "ansatz": "bloch", "vortex_number": 1.0}),
    ("Ferromagnetic superconductor", {"mode": "initial", "ansatz": "neel", "vortex_number": 2.0,
                                      "skyrmion_rotation": 0.3}),
    ("Spin triplet superconducting ferromagnet", {"mode": "ground"}),
    ("Spin triplet superconducting ferromagnet", {"mode": "initial", "ansatz": "anti"}),
    ("Bose-Einstein condensate", {"mode": "ground"}),
    ("Bose-Einstein condensate", {"mode": "initial", "vortex_number": 3.0}),
    ("Bose-Einstein condensate", {"mode": "initial", "vortex_number": -8.0}),
]


def _both(name, cfg, precision="fp64"):
    th = ssa.load_theory(name)
    sim = ssa.Simulation(_PARAMS[name](th), th, precision=precision)
    sim.initialize(cfg)
    F_host = sim.get_field()
    sim.set_velocity(np.full(sim.field_shape, 0.25))
    sim.initialize({**cfg, "device": True})
    return sim, F_host, sim.get_field(), sim.get_velocity()


@pytest.mark.parametrize("name,cfg", _CASES, ids=[f"{n}-{i}" for i, (n, _) in enumerate(_CASES)])
def test_device_init_matches_host(name, cfg):
    sim, F_host, F_dev, V_dev = _both(name, cfg)
    diff = np.abs(F_dev - F_host).max()
    scale = max(1.0, np.abs(F_host).max())
    print(f"[init {name} {cfg}] max|F|={np.abs(F_host).max():.3e} maxdiff={diff:.3e}", flush=True)
    assert diff <= 1e-11 * scale
    assert not np.any(V_dev)
    E_host = sim.compute_energy()
    sim.set_field(F_host)
    assert H.rel(sim.compute_energy(), E_host) < 1e-11
    sim.close()


@pytest.mark.parametrize("name,cfg", [("Chiral magnet", {"mode": "initial", "ansatz": "neel"}),
                                      ("Ginzburg-Landau superconductor", {"mode": "initial"}),
                                      ("Ferromagnetic superconductor", {"mode": "initial"})])
def test_device_init_fp32(name, cfg):
    sim, F_host, F_dev, _ = _both(name, cfg, precision="fp32")
    scale = max(1.0, np.abs(F_host).max())
    assert np.abs(F_dev - F_host).max() <= 1e-6 * scale
    sim.close()


def test_device_noise_unit_vector():
    """Gaussian noise on m, renormalised; Phi untouched; same seed -> same field."""
    th = ssa.load_theory("Chiral magnet")
    sim = ssa.Simulation(H.chiral_params(th, 64, 56), th)
    clean = {"mode": "initial", "ansatz": "bloch"}
    sim.initialize(clean)
    F0 = sim.get_field()
    noisy = {**clean, "noise": 0.05, "seed": 11}
    sim.initialize({**noisy, "device": True})
    F1 = sim.get_field()
    sim.initialize({**noisy, "device": True})
    assert np.array_equal(sim.get_field(), F1)
    sim.initialize({**noisy, "seed": 12, "device": True})
    assert not np.array_equal(sim.get_field(), F1)
    norm = np.sqrt((F1[:3] ** 2).sum(axis=0))
    assert np.abs(norm - 1.0).max() < 1e-12
    assert not np.any(F1[3])
    d = F1[:3] - F0[:3]
    assert 0.01 < d.std() < 0.06
    sim.close()


def test_device_noise_statistics():
    """ss_add_noise alone on a ground state: uniform U(-a, a) and Gaussian N(0, a^2)
    on the stated components only, with no renormalisation."""
    th = ssa.load_theory("Ginzburg-Landau superconductor")
    sim = ssa.Simulation(H.gl_params(th, 256, 200), th)
    L = _lib.lib()
    n = 256 * 200
    for gaussian, a in ((0, 0.3), (1, 0.2)):
        sim.initialize({"mode": "ground", "device": True})
        F0 = sim.get_field()
        assert L.ss_add_noise(sim._ctx, 2, 4, gaussian, a, 1234, 0, 0.0) == 0
        F1 = sim.get_field()
        assert np.array_equal(F1[:2], F0[:2])
        d = (F1[2:] - F0[2:]).ravel()
        assert abs(d.mean()) < 5.0 * a / np.sqrt(2 * n)
        if gaussian:
            assert abs(d.std() / a - 1.0) < 0.02
            assert np.abs(d).max() > 3.5 * a
        else:
            assert np.abs(d).max() <= a
            assert np.abs(d).max() > 0.999 * a
            assert abs(d.std() / (a / np.sqrt(3.0)) - 1.0) < 0.02
        # the two components draw independent values
        assert abs(np.corrcoef(d[:n], d[n:])[0, 1]) < 0.02
    sim.close()


def test_device_noise_gauge_theory_vortex():
    """GL vortex mode: uniform noise on Re/Im psi only, as the host generator adds it."""
    th = ssa.load_theory("Ginzburg-Landau superconductor")
    sim = ssa.Simulation(H.gl_params(th, 48, 40), th)
    cfg = {"mode": "vortex", "vortex_number": 2.0, "scale": 0.5}
    sim.initialize(cfg)
    F0 = sim.get_field()
    sim.initialize({**cfg, "noise": 0.02, "seed": 3, "device": True})
    F1 = sim.get_field()
    assert np.abs(F1[:2] - F0[:2]).max() <= 1e-11
    d = F1[2:] - F0[2:]
    assert 0.0 < np.abs(d).max() <= 0.02 + 1e-12
    sim.close()


def test_device_init_rejects_unavailable_modes():
    L = _lib.lib()
    th = ssa.load_theory("Chiral magnet")
    sim = ssa.Simulation(H.chiral_params(th, 32, 32), th)
    d = _lib.InitDesc()
    d.mode = _lib.INIT_VORTEX
    assert L.ss_init_field(sim._ctx, C.byref(d)) != 0
    assert b"not available" in L.ss_last_error(sim._ctx)
    d.mode = _lib.INIT_INITIAL
    d.ansatz = 5
    assert L.ss_init_field(sim._ctx, C.byref(d)) != 0
    d.ansatz = 0
    d.mode = _lib.INIT_SKYRMIONS
    d.nsol = 2
    assert L.ss_init_field(sim._ctx, C.byref(d)) != 0  # no soliton list
    assert L.ss_add_noise(sim._ctx, 0, 9, 0, 0.1, 1, 0, 0.0) != 0
    sim.close()
    th = ssa.load_theory("Bose-Einstein condensate")
    sim = ssa.Simulation(H.bec_params(th, 32, 32), th)
    d = _lib.InitDesc()
    d.mode = _lib.INIT_SKYRMIONS
    assert L.ss_init_field(sim._ctx, C.byref(d)) != 0
    sim.close()


_ORACLE_CASES = [c for c in _CASES
                 if c[0] in ("Chiral magnet", "Baby Skyrme model", "Ginzburg-Landau superconductor")
                 and c[1].get("mode") in ("ground", "initial", "skyrmions")
                 and "skyrmion_number" not in c[1]]


@pytest.mark.parametrize("name,cfg", _ORACLE_CASES, ids=[f"{n}-{c['mode']}-{i}" for i, (n, c) in enumerate(_ORACLE_CASES)])
def test_device_initial_conditions_match_the_oracle(name, cfg):
    """ss_init_field against oracle/initial.py, the restatement of the
    reference's initial-condition kernels (chiral_magnet/initial_config.py:
    16-386, baby_skyrme/initial_config.py:18-383,
    ginzburg_landau_superconductor/initial_config.py:14-313), not against the
    product's own host generators."""
    from tests.test_oracle_pinning_charges import oracle_initial_field
    th = ssa.load_theory(name)
    sim = ssa.Simulation(_PARAMS[name](th), th)
    try:
        sim.initialize({**cfg, "device": True})
        F_dev = sim.get_field()
        F_or = oracle_initial_field(name, cfg, sim.p_i_h, sim.p_f_h)
        err = np.abs(F_dev - F_or).max()
        print(f"[init vs oracle] {name} {cfg}: max|dF| = {err:.2e}", flush=True)
        assert err <= 1e-11 * max(1.0, np.abs(F_or).max())
    finally:
        sim.close()

