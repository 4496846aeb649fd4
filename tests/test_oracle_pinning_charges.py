"""Pins for the oracle's charge densities, DMI table, initial conditions and
its C++ restatement, each against a fact that does not come from retyping the
oracle's own formula:

* discrete Stokes / divergence identities on fields with a known flux or a
  polynomial form (the 4th-order stencils are exact on polynomials of degree
  <= 4, and a localised flux tube has a known total flux);
* the reference's own documentation of which DMI goes with which skyrmion
  type (chiral_magnet/params.py:6,72 and README.md:126: Dresselhaus with
  Bloch, Rashba with Neel, Heusler with anti-skyrmions) -- each DMI must
  lower the energy of its documented partner and leave the other two
  (almost) untouched;
* the Noether density as the A0-derivative of the energy density (the energy
  carries q^2 A0^2 |psi|^2 / 2 through A^2);
* topological facts of the initial conditions: unit norm, the charge of each
  ansatz, additivity of charges under stereographic composition, the winding
  of the vortex phase;
* the C++/OpenMP restatement against the NumPy oracle on one trajectory.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import core as ocore
from oracle import initial as oinit
from oracle.theories import anisotropic_superconductor as o_as
from oracle.theories import baby_skyrme as o_bs
from oracle.theories import chiral_magnet as o_cm
from oracle.theories import ferromagnetic_superconductor as o_fs
from oracle.theories import ginzburg_landau as o_gl
from oracle.theories import maxwell_chern_simons_higgs as o_mcsh
from oracle.theories import spin_triplet as o_st

import soliton_solver_amd as ssa
from tests import helpers as H


def _interior_masks(p_i):
    X, Y, h = int(p_i[0]), int(p_i[1]), int(p_i[2])
    inx = np.zeros((X, Y), dtype=bool)
    iny = np.zeros((X, Y), dtype=bool)
    inx[h:X - h, :] = True
    iny[:, h:Y - h] = True
    return inx, iny


# -- chiral magnet: magnetic charge and the DMI table ----------------------

def test_magnetic_charge_density_is_minus_divergence():
    """m1 = a x, m2 = b y: -(d_x m1 + d_y m2) gv = -(a + b) gv where both
    first derivatives exist, -a gv or -b gv where only one does (the 2-site
    bands of the stencils), 0 in the corners."""
    th = ssa.load_theory("Chiral magnet")
    p_i, p_f = H.packed(th, H.chiral_params(th, 30, 26))
    g = oinit.grid_coords(p_i, p_f)
    a, b = 0.37, -1.21
    F = np.zeros((4, int(p_i[0]), int(p_i[1])))
    F[0] = a * g[0] + 0.3
    F[1] = b * g[1] - 0.1
    F[2] = 0.5
    D1 = ocore.first_derivatives(F, p_i, p_f)
    rho = o_cm.magnetic_charge_density(F, D1, p_i, p_f)
    inx, iny = _interior_masks(p_i)
    expect = -(a * inx + b * iny) * float(p_f[4])
    assert np.allclose(rho, expect, rtol=1e-10, atol=1e-14)


def _dmi_energy(p_i, p_f, F, flag):
    """The DMI part of the energy: E(flag set) - E(no DMI flag)."""
    q = p_i.copy()
    q[11:15] = 0
    E0 = ocore.theory_energy(o_cm, F, q, p_f)
    if flag is not None:
        q[flag] = 1
    return ocore.theory_energy(o_cm, F, q, p_f) - E0


def test_dmi_table_favours_the_documented_skyrmion_type():
    th = ssa.load_theory("Chiral magnet")
    params = H.chiral_params(th, 72, 72, xsize=10.0, demag=False)
    p_i, p_f = H.packed(th, params)
    fields = {k: np.concatenate([oinit.centred_ansatz(p_i, p_f, k, 1.0), np.zeros((1, 72, 72))])
              for k in ("bloch", "neel", "anti")}
    partner = {11: "bloch", 12: "neel", 13: "anti"}  # dresselhaus, rashba, heusler flags of p_i
    table = {}
    for flag, kind in partner.items():
        e = {k: _dmi_energy(p_i, p_f, F, flag) for k, F in fields.items()}
        table[flag] = e
        scale = abs(e[kind])
        assert e[kind] < 0.0, (flag, e)
        for other in fields:
            if other != kind:
                # the other two ansatze average the DMI density to zero over the angle
                assert abs(e[other]) < 1e-2 * scale, (flag, other, e)
    # hybrid (p_i[14]) = (Dresselhaus + Rashba)/sqrt(2) entry by entry (chiral_magnet/kernels.py:180-183)
    for k, F in fields.items():
        eh = _dmi_energy(p_i, p_f, F, 14)
        assert abs(eh - (table[11][k] + table[12][k]) / math.sqrt(2.0)) < 1e-12 * max(1.0, abs(eh))


def test_dmi_energy_of_a_helix_is_linear_in_its_wavenumber():
    """m = (sin kx, 0, cos kx) (a cycloid in the x-z plane): m x d_x m has only
    a y component, k, so of the tables only Rashba's (j = y, i = x) entry D10 =
    +1 couples: the Rashba DMI energy per site is +k gv at every x-interior
    site, with opposite signs for the two chiralities, and Dresselhaus and
    Heusler (diagonal tables) give 0."""
    th = ssa.load_theory("Chiral magnet")
    p_i, p_f = H.packed(th, H.chiral_params(th, 40, 12, demag=False))
    g = oinit.grid_coords(p_i, p_f)
    for k in (0.9, -0.9):
        F = np.zeros((4, 40, 12))
        F[0] = np.sin(k * g[0])
        F[2] = np.cos(k * g[0])
        D1 = ocore.first_derivatives(F, p_i, p_f)
        q = p_i.copy()
        q[11:15] = 0
        base = o_cm.energy_density(F, D1, q, p_f)
        for flag, expect in ((12, 1.0), (11, 0.0), (13, 0.0)):
            r = q.copy()
            r[flag] = 1
            dmi = o_cm.energy_density(F, D1, r, p_f) - base
            inx, _ = _interior_masks(p_i)
            # the 4th-order stencil maps sin(kx), cos(kx) to k_h cos(kx), -k_h sin(kx) with
            # k_h = (4/3 sin(kh) - 1/6 sin(2kh))/h, so m x d_x m = (0, k_h, 0) exactly
            h = float(p_f[2])
            k_h = (4.0 / 3.0 * math.sin(k * h) - 1.0 / 6.0 * math.sin(2.0 * k * h)) / h
            assert np.allclose(dmi[inx], expect * k_h * float(p_f[4]), rtol=1e-11, atol=1e-14), (k, flag)
            assert np.all(np.abs(dmi[~inx]) < 1e-15)


# -- Maxwell-Chern-Simons-Higgs: electric and Noether charge ----------------

def test_electric_charge_density_is_minus_laplacian_of_a0():
    th = ssa.load_theory("Anyon superconductor")
    p_i, p_f = H.packed(th, H.mcsh_params(th, 28, 24))
    g = oinit.grid_coords(p_i, p_f)
    al, be = 0.31, -0.12
    F = np.zeros((5, 28, 24))
    F[4] = al * (g[0] - 1.3) ** 2 + be * (g[1] + 0.4) ** 2 + 0.2 * g[0] - 0.7
    D1 = ocore.first_derivatives(F, p_i, p_f)
    D2 = ocore.second_derivatives(F, p_i, p_f)
    rho = o_mcsh.electric_charge_density(F, D1, D2, p_i, p_f)
    inx, iny = _interior_masks(p_i)
    expect = -(2.0 * al * inx + 2.0 * be * iny) * float(p_f[4])
    assert np.allclose(rho, expect, rtol=1e-9, atol=1e-13)


def test_noether_density_is_the_a0_derivative_of_the_energy():
    """With a uniform A0 (no gradient), the energy density's A0 dependence is
    q^2 A0^2 |psi|^2 gv / 2, so the Noether density q^2 A0 |psi|^2 gv is its
    A0 derivative (central difference, exact for a quadratic); q != 1 so that
    a dropped factor of q shows."""
    th = ssa.load_theory("Anyon superconductor")
    p_i, p_f = H.packed(th, H.mcsh_params(th, 20, 18, q=1.7))
    rng = np.random.default_rng(3)
    F = rng.normal(size=(5, 20, 18))
    c, d = 0.45, 1e-3
    F[4] = c
    D1 = ocore.first_derivatives(F, p_i, p_f)
    Fp, Fm = F.copy(), F.copy()
    Fp[4] = c + d
    Fm[4] = c - d
    dE = (o_mcsh.energy_density(Fp, D1, p_i, p_f) - o_mcsh.energy_density(Fm, D1, p_i, p_f)) / (2 * d)
    rho = o_mcsh.noether_charge_density(F, D1, p_i, p_f)
    assert np.allclose(rho, dE, rtol=1e-8, atol=1e-12)


# -- vortex numbers: flux tubes of known flux --------------------------------

def _flux_tube(p_i, p_f, flux, w, comps=(0, 1)):
    """A = flux g(r) / (2 pi r) theta-hat with g = 1 - exp(-r^2/w^2), centred:
    B = d_x A2 - d_y A1 = flux g'(r) / (2 pi r), total flux = flux (Stokes)."""
    nf, X, Y = int(p_i[4]), int(p_i[0]), int(p_i[1])
    g = oinit.grid_coords(p_i, p_f)
    x = g[0] - g[0, X // 2, 0] - 0.013
    y = g[1] - g[1, 0, Y // 2] + 0.021
    r2 = x * x + y * y
    a_over_r = flux * (1.0 - np.exp(-r2 / (w * w))) / (2.0 * math.pi * r2)
    F = np.zeros((nf, X, Y))
    F[comps[0]] = -a_over_r * y
    F[comps[1]] = a_over_r * x
    return F


_VORTEX = [
    ("Ginzburg-Landau superconductor", o_gl, lambda th, q: H.gl_params(th, 96, 96, lsx=0.12, q=q), 6, (0, 1), {}),
    ("Anisotropic superconductor", o_as, lambda th, q: H.as_params(th, 96, 96, xsize=11.4, Q=q), 7, (0, 1), {}),
    ("Anyon superconductor", o_mcsh, lambda th, q: H.mcsh_params(th, 96, 96, xsize=11.4, q=q), 6, (0, 1), {}),
    ("Ferromagnetic superconductor", o_fs, lambda th, q: H.fs_params(th, 96, 96, xsize=11.4, q=q), 6, (3, 4),
     {"which": 2}),
    ("Spin triplet superconducting ferromagnet", o_st, lambda th, q: H.st_params(th, 96, 96, xsize=11.4, q=q), 6,
     (3, 4), {"which": 2}),
]


@pytest.mark.parametrize("name,o,mkp,qidx,comps,kw", _VORTEX, ids=[c[0] for c in _VORTEX])
@pytest.mark.parametrize("q,N", [(1.0, 1), (1.0, -2), (2.0, 1)])
def test_vortex_number_of_a_flux_tube(name, o, mkp, qidx, comps, kw, q, N):
    """A winding-N vortex in the coupling q carries flux 2 pi N / q; the
    reference's density (flux density x gv / (2 pi q)) sums to N / q^2
    (N itself at q = 1)."""
    th = ssa.load_theory(name)
    p_i, p_f = H.packed(th, mkp(th, q))
    assert float(p_f[qidx]) == q
    F = _flux_tube(p_i, p_f, 2.0 * math.pi * N / q, w=1.2, comps=comps)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    # interior sites only: in the 2-site bands one of the two derivatives is 0
    # (the stencils' masks), which leaves the pure-gauge tail's other half there
    inx, iny = _interior_masks(p_i)
    total = float(np.sum(o.vortex_density(F, D1, p_i, p_f, **kw)[inx & iny]))
    assert abs(total - N / q ** 2) < 1e-6 * abs(N), (total, N / q ** 2)


@pytest.mark.parametrize("o", [o_fs, o_st], ids=["ferromagnetic", "spin_triplet"])
def test_in_plane_flux_components_are_the_curl_of_a3(o):
    """which 0 is B_x = d_y A3 and which 1 is B_y = -d_x A3 (B = curl A with
    A3 independent of z), checked on a quadratic A3 at interior sites."""
    th = ssa.load_theory("Ferromagnetic superconductor" if o is o_fs else "Spin triplet superconducting ferromagnet")
    p_i, p_f = H.packed(th, (H.fs_params if o is o_fs else H.st_params)(th, 26, 22, q=1.3))
    g = oinit.grid_coords(p_i, p_f)
    F = np.zeros((int(p_i[4]), 26, 22))
    F[5] = 0.3 * g[0] ** 2 - 0.2 * g[0] * g[1] + 0.7 * g[1]
    D1 = ocore.first_derivatives(F, p_i, p_f)
    inx, iny = _interior_masks(p_i)
    s = float(p_f[4]) / (2.0 * math.pi * 1.3)
    bx = o.vortex_density(F, D1, p_i, p_f, which=0)
    by = o.vortex_density(F, D1, p_i, p_f, which=1)
    assert np.allclose(bx[iny], ((-0.2 * g[0] + 0.7) * s)[iny], rtol=1e-9, atol=1e-13)
    assert np.allclose(by[inx], (-(0.6 * g[0] - 0.2 * g[1]) * s)[inx], rtol=1e-9, atol=1e-13)
    assert np.all(bx[~iny] == 0.0) and np.all(by[~inx] == 0.0)


# -- initial conditions -------------------------------------------------------

def test_initial_profile_and_angle():
    assert oinit.profile_magnetization(0.0, 2.0) == math.pi
    assert abs(oinit.profile_magnetization(2.0, 2.0) - math.pi * math.exp(-4.0)) < 1e-15
    assert oinit.profile_magnetization(2.0 + 1e-9, 2.0) == 0.0
    assert oinit.angle(0.0, 0.0, 0.0, 0.3) == -0.3
    assert abs(oinit.angle(1.0, -1e-9, 1.0, 0.0) - (2.0 * math.pi - 1e-9)) < 1e-12
    assert abs(oinit.angle(-1.0, -1e-9, 1.0, 0.0) - (math.pi + 1e-9)) < 1e-12
    assert abs(oinit.angle(0.0, 1.0, 1.0, 0.25) - (math.pi / 2 - 0.25)) < 1e-15
    assert oinit.profile_superconductor(3.0, 2.0, 0.2) == 1.0
    assert abs(oinit.profile_superconductor(1.0, 2.0, 0.2) - math.tanh(0.2)) < 1e-16


def _skyrmion_number(p_i, p_f, m):
    F = np.concatenate([m, np.zeros((1,) + m.shape[1:])])
    D1 = ocore.first_derivatives(F, p_i, p_f)
    return float(np.sum(o_cm.skyrmion_density(F, D1, p_i, p_f)))


@pytest.mark.parametrize("kind,Q", [("bloch", -1.0), ("neel", -1.0), ("anti", 1.0)])
def test_centred_ansatz_is_a_unit_field_of_charge_one(kind, Q):
    th = ssa.load_theory("Chiral magnet")
    p_i, p_f = H.packed(th, H.chiral_params(th, 121, 121, xsize=10.0))
    m = oinit.centred_ansatz(p_i, p_f, kind, 1.0, rotation=0.4)
    assert np.allclose((m ** 2).sum(0), 1.0, atol=1e-15)
    assert abs(m[2, 60, 60] + 1.0) < 1e-12  # the profile angle is pi at the centre
    assert np.allclose(m[:, 0, 0], [0.0, 0.0, 1.0])  # and 0 beyond its radius
    # (the profile is cut off at its radius, where m3 jumps from cos(pi e^-4) to 1)
    assert abs(_skyrmion_number(p_i, p_f, m) - Q) < 5e-3


def test_stereographic_composition_adds_charges():
    th = ssa.load_theory("Chiral magnet")
    p_i, p_f = H.packed(th, H.chiral_params(th, 140, 100, xsize=14.0))
    X, Y = 140, 100
    ground = np.zeros((3, X, Y))
    ground[2] = 1.0
    # ground + one soliton = that soliton (w of the ground state is 0)
    one = oinit.compose_skyrmion(ground.copy(), p_i, p_f, 7.0, 5.0, 1.8, "anti", 1.0)
    g = oinit.grid_coords(p_i, p_f)
    xc_site = np.argmin(np.abs(g[0][:, 0] - 7.0))
    assert abs(one[2, xc_site, 50] + 1.0) < 1e-2
    assert abs(_skyrmion_number(p_i, p_f, one) - 1.0) < 5e-3
    two = oinit.compose_skyrmion(one.copy(), p_i, p_f, 3.0, 5.0, 1.8, "anti", 1.0)
    three = oinit.compose_skyrmion(two.copy(), p_i, p_f, 11.0, 5.0, 1.8, "bloch", 1.0)
    assert np.allclose((three ** 2).sum(0), 1.0, atol=1e-14)
    assert abs(_skyrmion_number(p_i, p_f, two) - 2.0) < 1e-2
    assert abs(_skyrmion_number(p_i, p_f, three) - 1.0) < 1e-2


def test_gl_vortex_phase_winds_n_times():
    th = ssa.load_theory("Ginzburg-Landau superconductor")
    p_i, p_f = H.packed(th, H.gl_params(th, 61, 61, lsx=0.25))
    for N in (1.0, 2.0, -1.0):
        F = oinit.gl_centred_vortex(p_i, p_f, N, ainf=0.5)
        # phase increments around the square ring at distance 20 sites from the centre
        ring = [(30 + i, 10) for i in range(-20, 20)] + [(50, 10 + j) for j in range(40)] + \
               [(50 - i, 50) for i in range(40)] + [(10, 50 - j) for j in range(40)]
        ph = np.array([math.atan2(F[3, x, y], F[2, x, y]) for x, y in ring + ring[:1]])
        d = np.angle(np.exp(1j * np.diff(ph)))
        assert abs(d.sum() / (2 * math.pi) - N) < 1e-12
        u1 = float(p_f[8])
        r = 20 * 0.25
        assert abs(math.hypot(F[2, 50, 30], F[3, 50, 30]) - u1 * math.tanh(0.2 * r)) < 1e-12
    ground = oinit.gl_ground(p_i, p_f)
    comp = oinit.gl_compose_vortex(ground.copy(), p_i, p_f, 7.5, 7.5)
    assert np.allclose(comp[2] ** 2 + comp[3] ** 2,
                       (float(p_f[8]) * np.tanh(0.2 * np.hypot(*(oinit.grid_coords(p_i, p_f) - 7.5)))) ** 2)


_HOST_CASES = [
    ("Chiral magnet", lambda th: H.chiral_params(th, 48, 40), {"mode": "initial", "ansatz": "bloch"}),
    ("Chiral magnet", lambda th: H.chiral_params(th, 48, 40),
     {"mode": "initial", "ansatz": "neel", "skyrmion_rotation": 0.4}),
    ("Chiral magnet", lambda th: H.chiral_params(th, 48, 40),
     {"mode": "skyrmions", "ansatz": "anti", "skyrmions": [(2.6, 2.3, 0.9), (7.1, 5.9, 1.1)]}),
    ("Baby Skyrme model", lambda th: H.bs_params(th, 52, 44, skyrmion_number=2.0),
     {"mode": "initial", "ansatz": "neel"}),
    ("Baby Skyrme model", lambda th: H.bs_params(th, 52, 44, skyrmion_number=0.0),
     {"mode": "skyrmions", "skyrmion_rotation": 0.2, "skyrmions": [(6.0, 7.5, 2.0), (13.3, 9.1, 2.5)]}),
    ("Ginzburg-Landau superconductor", lambda th: H.gl_params(th, 48, 40), {"mode": "ground"}),
    ("Ginzburg-Landau superconductor", lambda th: H.gl_params(th, 48, 40), {"mode": "initial", "vortex_number": 2.0}),
]


def oracle_initial_field(name, cfg, p_i, p_f):
    """The oracle's field for a product configuration dict (modes the
    reference has: ground, centred ansatz / vortex, composed skyrmions with
    the product's per-soliton radius R as profile radius 2R)."""
    X, Y = int(p_i[0]), int(p_i[1])
    mode = cfg.get("mode", "initial")
    if name == "Ginzburg-Landau superconductor":
        if mode == "ground":
            return oinit.gl_ground(p_i, p_f)
        return oinit.gl_centred_vortex(p_i, p_f, float(cfg.get("vortex_number", 1.0)), float(p_f[10]))
    skN = float(p_f[9] if name == "Chiral magnet" else p_f[8])
    kind = cfg.get("ansatz", "bloch")
    rot = float(cfg.get("skyrmion_rotation", 0.0))
    nf = int(p_i[4])
    F = np.zeros((nf, X, Y))
    if mode == "ground":
        F[2] = 1.0
    elif mode == "skyrmions":
        m = np.zeros((3, X, Y))
        m[2] = 1.0
        for xc, yc, R in cfg["skyrmions"]:
            oinit.compose_skyrmion(m, p_i, p_f, xc, yc, 2.0 * R, kind, skN, rot)
        F[:3] = m
    else:
        F[:3] = oinit.centred_ansatz(p_i, p_f, kind, skN, rot)
    return F


@pytest.mark.parametrize("name,mkp,cfg", _HOST_CASES, ids=[f"{c[0]}-{c[2]['mode']}-{i}" for i, c in enumerate(_HOST_CASES)])
def test_host_generators_match_the_oracle(name, mkp, cfg):
    th = ssa.load_theory(name)
    p_i, p_f = H.packed(th, mkp(th))
    F = np.zeros((int(p_i[4]), int(p_i[0]), int(p_i[1])))
    V = np.zeros_like(F)
    th.initial_config.initialize(Field=F, Velocity=V, grid=H.grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f, config=cfg)
    Fo = oracle_initial_field(name, cfg, p_i, p_f)
    assert np.abs(F - Fo).max() < 1e-13, np.abs(F - Fo).max()


# -- the C++ restatement ------------------------------------------------------

def test_cpu_restatement_matches_numpy_oracle_trajectory():
    """oracle/cpu_ref.cpp (the bench's CPU baseline) against the NumPy oracle:
    30 steps of the chiral magnet (Heusler, demag, unit magnetisation) on a
    40 x 52 lattice from one seeded field, with a forced arrest on step 1."""
    from oracle import cpu_build
    th = ssa.load_theory("Chiral magnet")
    p_i, p_f = H.packed(th, H.chiral_params(th, 40, 52))
    F0 = H.chiral_field(th, p_i, p_f, seed=5)
    V0 = np.zeros_like(F0)
    Fn, Vn = F0.copy(), V0.copy()
    En, errn = ocore.relax(o_cm, Fn, Vn, p_i, p_f, 30, prev_energy=-np.inf)
    Fc, Vc = np.ascontiguousarray(F0.copy()), np.ascontiguousarray(V0.copy())
    Ec, errc = cpu_build.chiral_relax(p_i, p_f, Fc, Vc, -np.inf, 30, nthreads=2)
    assert np.allclose(Ec, En, rtol=1e-13, atol=0)
    assert np.allclose(errc, errn, rtol=1e-10, atol=1e-13)
    assert np.abs(Fc - Fn).max() < 1e-13
    assert np.abs(Vc - Vn).max() < 1e-12
