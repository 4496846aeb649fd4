"""The compensated FP32 form ("fp32c", SS_PREC_FP32C) against the FP64 form.

The form keeps FP32 storage and FP32 stencils (l_k, V and the stage fields as
in "fp32") but holds the field as two FP32 words, hi + lo: pass 4 adds each
step's increment to hi + lo in FP64 and splits the sum again, and pass 1 takes
the energy (the input of the arrest test) of hi + lo in FP64.  What is pinned:

* the split is exact to 2^-48: FP64 -> fp32c -> FP64 returns the field within
  4e-15 max|F|, and the energy of the split field equals the FP64 form's to
  1e-12 (plain FP32 storage moves it by ~1e-7);
* per theory, 20 steps from the same field stay within the FP32 form's
  tolerances of the FP64 form (1e-5 relative energy);
* with a time step so small that every per-step increment is below an FP32 ulp
  of the field, the displacement after 400 steps matches the FP64 form's to
  1e-3 of its size, where plain FP32 storage loses most of it;
* autotuning (which times the in-place pass 4 on a copy of the low words)
  leaves the state unchanged; the form is refused where it does not exist.
"""
from __future__ import annotations

import numpy as np
import pytest

import soliton_solver_amd as ssa
from soliton_solver_amd import _lib
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _sim(theory, params, F, precision="fp64"):
    sim = ssa.Simulation(params, theory, precision=precision)
    sim.initialize({"mode": "array", "field": F})
    return sim


def _chiral(n=64, m=56, seed=4, **kw):
    th = ssa.load_theory("Chiral magnet")
    p = H.chiral_params(th, n, m, **kw)
    p_i, p_f = H.packed(th, p)
    return th, p, H.chiral_field(th, p_i, p_f, seed=seed)


def test_split_round_trip_and_energy():
    th, p, F = _chiral()
    a = _sim(th, p, F)
    E64 = a.compute_energy()
    c = _sim(th, p, F, precision="fp32c")
    Fc = c.get_field()
    scale = np.abs(F).max()
    print(f"[fp32c split] max|F_c - F|={np.abs(Fc - F).max():.2e} relE={H.rel(c.compute_energy(), E64):.2e}")
    assert np.abs(Fc - F).max() <= 4e-15 * scale
    assert H.rel(c.compute_energy(), E64) < 1e-12
    # the same through set_precision on the device
    a.set_precision("fp32c")
    assert H.rel(a.compute_energy(), E64) < 1e-12
    a.set_precision("fp64")
    assert np.abs(a.get_field() - F).max() <= 4e-15 * scale
    # plain FP32 storage rounds the field: its energy moves at the FP32 level
    b = _sim(th, p, F, precision="fp32")
    print(f"[fp32 storage] relE={H.rel(b.compute_energy(), E64):.2e}")
    assert H.rel(b.compute_energy(), E64) < 1e-6
    # fp32 -> fp32c starts the low words at zero (the field is the FP32 one);
    # fp32c -> fp32 rounds hi + lo to FP32
    Fb = b.get_field()
    b.set_precision("fp32c")
    assert np.array_equal(b.get_field(), Fb)
    c.set_precision("fp32")
    assert np.abs(c.get_field() - F).max() <= 6e-8 * scale
    for s in (a, b, c):
        s.close()


_CASES = {
    "Chiral magnet": (lambda th: H.chiral_params(th, 60, 52), lambda th, pi, pf: H.chiral_field(th, pi, pf, seed=3)),
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


@pytest.mark.parametrize("variants", [None, [3, 3, 3, 0], [6, 6, 6, 7]])
@pytest.mark.parametrize("name", list(_CASES))
def test_fp32c_tracks_fp64_for_every_theory(name, variants):
    """20 steps from the same field, in the autotune-free default shapes, in
    one-wavefront strips with a tile pass 4, and in two-wavefront strips."""
    th = ssa.load_theory(name)
    mkp, mkf = _CASES[name]
    p = mkp(th)
    p_i, p_f = H.packed(th, p)
    F = mkf(th, p_i, p_f)
    a = _sim(th, p, F)
    c = _sim(th, p, F, precision="fp32c")
    if variants is not None:
        c.set_tile_variants(variants)
    E0 = a.compute_energy()
    assert H.rel(c.compute_energy(), E0) < 1e-12
    Ea, _, _ = a.steps(E0, 20)
    Ec, _, _ = c.steps(E0, 20)
    dF = np.abs(a.get_field() - c.get_field()).max() / max(1.0, np.abs(F).max())
    print(f"[fp32c {name} {variants}] after 20 steps relE={H.rel(Ec, Ea):.2e} max|dF|={dF:.2e}")
    assert H.rel(Ec, Ea) < 1e-5
    assert dF < 1e-3
    # the energy the step returns is the compensated one
    assert H.rel(Ec, c.compute_energy()) < 1e-12
    a.close()
    c.close()


@pytest.mark.parametrize("name", list(_CASES))
def test_fp32c_paired_matches_unpaired(name):
    """The compensated form's paired schedule (passes 1+2 in one sweep that
    carries the low words, V read before the arrest decision, where that ring
    fits, else pass 1 and pass 2 after the decision; then passes 3+4 with the
    compensated combine in one sweep) against its 2+3 / 4 schedule in the same
    strips: 20 steps in calls of 7 and 13, then 2 steps from prev_energy = -inf
    (an arrested first step)."""
    th = ssa.load_theory(name)
    mkp, mkf = _CASES[name]
    p = mkp(th)
    p_i, p_f = H.packed(th, p)
    F = mkf(th, p_i, p_f)
    a = _sim(th, p, F, precision="fp32c")
    b = _sim(th, p, F, precision="fp32c")
    for s in (a, b):
        s.set_tile_variants([6, 6, 6, 6])
    b.set_fused_stages(True, False)
    try:
        a.set_fused_stages(False, False, pairs=True)
    except ValueError:  # the two-field ring does not fit this theory
        a.close()
        b.close()
        pytest.skip(f"no paired sweeps for {name}")
    comp12 = a.tile_variants()["comp12"]
    if name == "Chiral magnet":  # the three-part ring of the compensated 1+2 sweep fits
        assert comp12
    E0 = a.compute_energy()
    Ea, Eb = E0, E0
    for n in (7, 13):
        Ea, _, _ = a.steps(Ea, n)
        Eb, _, _ = b.steps(Eb, n)
    scale = max(1.0, np.abs(F).max())
    dF = np.abs(a.get_field() - b.get_field()).max() / scale
    print(f"[fp32c paired {name}, 1+2 sweep {comp12}] relE={H.rel(Ea, Eb):.2e} max|dF|={dF:.2e}")
    assert H.rel(Ea, Eb) < 1e-7 and dF < 1e-5
    Ea, _, _ = a.steps(-np.inf, 2)
    Eb, _, _ = b.steps(-np.inf, 2)
    dF = np.abs(a.get_field() - b.get_field()).max() / scale
    dV = np.abs(a.get_velocity() - b.get_velocity()).max() / scale
    print(f"[fp32c paired {name} arrested] relE={H.rel(Ea, Eb):.2e} max|dF|={dF:.2e} max|dV|={dV:.2e}")
    assert H.rel(Ea, Eb) < 1e-7 and dF < 1e-5 and dV < 1e-5
    # the returned energy is the compensated one
    assert H.rel(Ea, a.compute_energy()) < 1e-12
    a.close()
    b.close()


def test_sub_ulp_increments_accumulate():
    """dt = 2e-6: every increment of the field per step is below an FP32 ulp of
    its O(1) components, so plain FP32 storage rounds most of them away while
    the compensated form keeps them."""
    th, p, F = _chiral(48, 40, seed=2, demag=False, time_step=2e-6)
    disp = {}
    for prec in ("fp64", "fp32c", "fp32"):
        s = _sim(th, p, F, precision=prec)
        F0 = s.get_field()
        E = s.compute_energy()
        E, _, _ = s.steps(E, 400)
        disp[prec] = s.get_field() - F0
        s.close()
    d64 = disp["fp64"]
    size = np.abs(d64).max()
    errc = np.abs(disp["fp32c"] - d64).max() / size
    err32 = np.abs(disp["fp32"] - d64).max() / size
    print(f"[sub-ulp] max|dF64|={size:.2e} fp32c err={errc:.2e} fp32 err={err32:.2e}")
    assert 1e-6 < size < 1e-4
    assert errc < 1e-3
    assert err32 > 10 * errc


def test_autotune_keeps_state():
    th, p, F = _chiral(96, 80)
    c = _sim(th, p, F, precision="fp32c")
    E = c.compute_energy()
    E, _, _ = c.steps(E, 10)
    F1, V1 = c.get_field(), c.get_velocity()
    info = c.autotune_tiles(reps=2)
    assert not info["fused41"]
    assert info["variants"][0] >= 3  # the compensated pass 1 streams
    assert np.array_equal(c.get_field(), F1)
    assert np.array_equal(c.get_velocity(), V1)
    c.close()


def test_fp32c_refused_where_unavailable():
    th = ssa.load_theory("Bose-Einstein condensate")
    with pytest.raises(ValueError):
        ssa.Simulation(H.bec_params(th, 32, 32), th, precision="fp32c")
    s = ssa.Simulation(H.bec_params(th, 32, 32), th)
    with pytest.raises(ValueError):
        s.set_precision("fp32c")
    s.close()
    th, p, F = _chiral(48, 40)
    c = _sim(th, p, F, precision="fp32c")
    with pytest.raises(ValueError):
        c.set_fused_stages(True, True)  # no 4+1 sweep in the compensated form
    c.close()
    with pytest.raises(ValueError):
        ssa.Simulation(p, th, precision="fp32c", variation="exact_adjoint")
    assert _lib.PREC_FP32C == 2


def test_reported_variants_are_the_ones_that_run():
    """Where a shape variant lacks a sweep (the compensated pass 1 streams; an
    FP64 chiral tile of 16 rows cannot hold two fields), the pass runs in the
    first strip variant and the context reports that variant."""
    th, p, F = _chiral(96, 80)
    c = _sim(th, p, F, precision="fp32c")
    assert c.tile_variants()["variants"][0] >= 3
    c.set_tile_variants([0, 0, 0, 0])
    assert c.tile_variants()["variants"] == [3, 0, 0, 0]
    c.close()
    a = _sim(th, p, F)
    a.set_fused_stages(True, True)
    a.set_tile_variants([0, 0, 0, 0])
    assert a.tile_variants()["variants"] == [0, 3, 0, 3]
    a.set_tile_variants([2, 2, 2, 2])  # 4-row tiles hold both fields
    assert a.tile_variants()["variants"] == [2, 2, 2, 2]
    a.close()


def test_minimize_fast_reaches_the_fp64_minimum():
    """Simulation.minimize_fast (compensated FP32 chunks, FP64 polish checks)
    against the plain FP64 relaxation from the same field, both to max|EG|
    < 1e-6: energies within 1e-7 relative (the precision form's bar), the
    same skyrmion number, and the context ends in FP64."""
    th, p, F = _chiral(96, 80, seed=6, demag=False)
    a = _sim(th, p, F)
    Ea, eaa, na = a.minimize(tol=1e-6, max_steps=300000, check_every=500)
    b = _sim(th, p, F)
    Eb, eab, n32, n64 = b.minimize_fast(tol=1e-6, max_steps=300000, chunk=4000)
    Qa, Qb = a.observables()["skyrmion_number"], b.observables()["skyrmion_number"]
    print(f"[minimize_fast] fp64: {na} steps E={Ea:.12f} err={eaa:.2e}; fast: {n32} fp32c + {n64} fp64 steps "
          f"E={Eb:.12f} err={eab:.2e} relE={H.rel(Eb, Ea):.2e} Q {Qa:.6f} {Qb:.6f}")
    assert eaa < 1e-6 and eab < 1e-6
    assert b.precision == "fp64"
    assert H.rel(Eb, Ea) < 1e-7
    assert round(Qa) == round(Qb)
    a.close()
    b.close()
