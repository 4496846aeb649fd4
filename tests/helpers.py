"""Shared builders for the tests: parameters, seeded fields, oracle runs."""
from __future__ import annotations

import numpy as np

from oracle import core as ocore
from oracle.theories import anisotropic_superconductor as o_as
from oracle.theories import baby_skyrme as o_bs
from oracle.theories import bose_einstein_condensate as o_bec
from oracle.theories import chiral_magnet as o_cm
from oracle.theories import ferromagnetic_superconductor as o_fs
from oracle.theories import maxwell_chern_simons_higgs as o_mcsh
from oracle.theories import ginzburg_landau as o_gl
from oracle.theories import liquid_crystal as o_lc
from oracle.theories import spin_triplet as o_st

ORACLES = {
    "Chiral magnet": o_cm,
    "Ginzburg-Landau superconductor": o_gl,
    "Baby Skyrme model": o_bs,
    "Anisotropic superconductor": o_as,
    "Bose-Einstein condensate": o_bec,
    "Anyon superconductor": o_mcsh,
    "Ferromagnetic superconductor": o_fs,
    "Spin triplet superconducting ferromagnet": o_st,
    "Liquid crystal": o_lc,
}


def chiral_params(theory, n=64, m=None, xsize=10.0, **kw):
    m = m or n
    base = dict(xlen=n, ylen=m, xsize=xsize, ysize=xsize * (m - 1) / (n - 1),
                J=40e-12, K=0.8e+6, D=4e-3, M=580e+3, B=0e-3, dmi_term="Heusler", demag=True,
                unit_magnetization=True)
    base.update(kw)
    return theory.params.default_params(**base)


def gl_params(theory, n=48, m=None, lsx=0.2, **kw):
    m = m or n
    base = dict(xlen=n, ylen=m, xsize=lsx * (n - 1), ysize=lsx * (m - 1), q=1.0, Lambda=1.0, u1=1.0,
                vortex_number=1.0, time_step=0.5 * lsx)
    base.update(kw)
    return theory.params.default_params(**base)


def packed(theory, params):
    p_i, p_f = theory.params.pack_device_params(params.resolved())
    return np.ascontiguousarray(p_i, dtype=np.int32), np.ascontiguousarray(p_f, dtype=np.float64)


def grid_of(p_i, p_f):
    X, Y = int(p_i[0]), int(p_i[1])
    g = np.empty((2, X, Y))
    g[0] = p_f[2] * np.arange(X)[:, None]
    g[1] = p_f[3] * np.arange(Y)[None, :]
    return g


def chiral_field(theory, p_i, p_f, seed=0, kind="anti", sigma=1e-3, radius_scale=1.0, rough=False):
    """Anti-skyrmion at the centre plus seeded Gaussian noise, unit norm; psi seeded small."""
    nf, X, Y = int(p_i[4]), int(p_i[0]), int(p_i[1])
    F = np.zeros((nf, X, Y))
    V = np.zeros_like(F)
    cfg = {"mode": "initial", "ansatz": kind, "noise": sigma, "seed": seed}
    theory.initial_config.initialize(Field=F, Velocity=V, grid=grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                     config=cfg)
    rng = np.random.default_rng(seed + 1000)
    if rough:
        F[:3] += rng.normal(0.0, 0.3, size=F[:3].shape)
        s = np.sqrt((F[:3] ** 2).sum(0))
        F[:3] /= s
    F[3] = 0.01 * rng.normal(size=F[3].shape)
    return F


def gl_field(theory, p_i, p_f, seed=0, amp=1e-3, velocity=False):
    nf, X, Y = int(p_i[4]), int(p_i[0]), int(p_i[1])
    F = np.zeros((nf, X, Y))
    V = np.zeros_like(F)
    theory.initial_config.initialize(Field=F, Velocity=V, grid=grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                     config={"mode": "vortex", "noise": amp, "seed": seed})
    rng = np.random.default_rng(seed + 7)
    F[0] += 0.05 * rng.normal(size=F[0].shape)
    F[1] += 0.05 * rng.normal(size=F[1].shape)
    if velocity:
        V = 0.01 * rng.normal(size=F.shape)
        return F, V
    return F


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def bs_params(theory, n=48, m=None, xsize=20.0, **kw):
    m = m or n
    h = xsize / (n - 1)
    base = dict(xlen=n, ylen=m, xsize=xsize, ysize=h * (m - 1), mpi=1.0, kappa=1.0, unit_magnetization=True,
                time_step=0.0957 * h)
    base.update(kw)
    return theory.params.default_params(**base)


def bs_field(theory, p_i, p_f, seed=0, k=1, amp=1e-3, kind="neel"):
    nf, X, Y = int(p_i[4]), int(p_i[0]), int(p_i[1])
    F = np.zeros((nf, X, Y))
    V = np.zeros_like(F)
    theory.initial_config.initialize(Field=F, Velocity=V, grid=grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                     config={"mode": "initial", "ansatz": kind, "skyrmion_number": k,
                                             "noise": amp, "seed": seed})
    return F


def as_betas(seed, alpha1=-1.0, alpha2=-0.5):
    """beta1..beta4 ~ U[0.2, 1] from the seed, redrawn until the quartic form
    beta1 s1^2 + beta2 s2^2 + (beta3 - 2 beta4) s1 s2 is bounded below (with
    beta4 > 0 the phase locks at theta = pi/2) and both vacuum values are real
    and non-zero; otherwise the energy is unbounded below and no relaxation
    converges."""
    rng = np.random.default_rng(seed)
    while True:
        b1, b2, b3, b4 = rng.uniform(0.2, 1.0, size=4)
        b3t = b3 - 2.0 * b4
        den = b3t * b3t - 4.0 * b1 * b2
        n1 = 2.0 * b2 * alpha1 - b3t * alpha2
        n2 = 2.0 * b1 * alpha2 - b3t * alpha1
        if b3t > -2.0 * np.sqrt(b1 * b2) + 0.1 and den != 0 and n1 / den > 0 and n2 / den > 0:
            return b1, b2, b3, b4


def as_params(theory, n=48, m=None, xsize=25.0, seed=0, **kw):
    """Config-5 style: alpha1 = -1, alpha2 = -0.5, beta1..beta4 ~ U[0.2, 1] from the seed,
    gamma1 = gamma2 = 1, gamma12 = 0.3."""
    m = m or n
    h = xsize / (n - 1)
    b = as_betas(seed)
    base = dict(xlen=n, ylen=m, xsize=xsize, ysize=h * (m - 1), alpha1=-1.0, alpha2=-0.5, beta1=b[0], beta2=b[1],
                beta3=b[2], beta4=b[3], gamma1=1.0, gamma2=1.0, gamma12=0.3, Q=1.0, kappa=0.673274)
    base.update(kw)
    return theory.params.default_params(**base)


def as_field(theory, p_i, p_f, seed=0, amp=1e-2, mode="initial"):
    nf, X, Y = int(p_i[4]), int(p_i[0]), int(p_i[1])
    F = np.zeros((nf, X, Y))
    V = np.zeros_like(F)
    theory.initial_config.initialize(Field=F, Velocity=V, grid=grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                     config={"mode": mode, "noise": amp, "seed": seed})
    return F


def bec_params(theory, n=48, m=None, xsize=12.0, **kw):
    m = m or n
    h = xsize / (n - 1)
    base = dict(xlen=n, ylen=m, xsize=xsize, ysize=h * (m - 1), beta=50.0, omega_rot=0.5, vortex_number=2.0)
    base.update(kw)
    return theory.params.default_params(**base)


def bec_field(theory, p_i, p_f, seed=0, amp=1e-3):
    nf, X, Y = int(p_i[4]), int(p_i[0]), int(p_i[1])
    F = np.zeros((nf, X, Y))
    V = np.zeros_like(F)
    theory.initial_config.initialize(Field=F, Velocity=V, grid=grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                     config={"mode": "ground", "noise": amp, "seed": seed})
    return F


def mcsh_params(theory, n=48, m=None, xsize=12.0, **kw):
    m = m or n
    h = xsize / (n - 1)
    base = dict(xlen=n, ylen=m, xsize=xsize, ysize=h * (m - 1), q=1.0, Lambda=1.5, kappa=0.7, u1=1.0)
    base.update(kw)
    return theory.params.default_params(**base)


def mcsh_field(theory, p_i, p_f, seed=0, amp=1e-2):
    nf, X, Y = int(p_i[4]), int(p_i[0]), int(p_i[1])
    F = np.zeros((nf, X, Y))
    V = np.zeros_like(F)
    theory.initial_config.initialize(Field=F, Velocity=V, grid=grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                     config={"mode": "initial", "noise": amp, "seed": seed})
    g = grid_of(p_i, p_f)
    F[4] = 0.1 * np.exp(-((g[0] - g[0].mean()) ** 2 + (g[1] - g[1].mean()) ** 2) / 4.0)
    return F


def fs_params(theory, n=48, m=None, xsize=16.0, **kw):
    """eta1 = -0.1 (eta2 derived = -0.0125), giving M0 = 1.369, u1 = 2.092."""
    m = m or n
    h = xsize / (n - 1)
    base = dict(xlen=n, ylen=m, xsize=xsize, ysize=h * (m - 1), q=0.8, eta1=-0.1, gamma=1.2,
                unit_magnetization=True)
    base.update(kw)
    return theory.params.default_params(**base)


def fs_field(theory, p_i, p_f, seed=0, amp=1e-2, ansatz="bloch", vortex_number=1.0):
    nf, X, Y = int(p_i[4]), int(p_i[0]), int(p_i[1])
    F = np.zeros((nf, X, Y))
    V = np.zeros_like(F)
    theory.initial_config.initialize(Field=F, Velocity=V, grid=grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                     config={"mode": "initial", "ansatz": ansatz, "vortex_number": vortex_number,
                                             "noise": amp, "seed": seed})
    return F


def st_params(theory, n=48, m=None, xsize=16.0, **kw):
    """Defaults give M0 = 1, u1 = u2 = sqrt(5/4)."""
    m = m or n
    h = xsize / (n - 1)
    base = dict(xlen=n, ylen=m, xsize=xsize, ysize=h * (m - 1), q=0.9, gamma=1.1, unit_magnetization=True)
    base.update(kw)
    return theory.params.default_params(**base)


def st_field(theory, p_i, p_f, seed=0, amp=1e-2, ansatz="bloch"):
    nf, X, Y = int(p_i[4]), int(p_i[0]), int(p_i[1])
    F = np.zeros((nf, X, Y))
    V = np.zeros_like(F)
    theory.initial_config.initialize(Field=F, Velocity=V, grid=grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                     config={"mode": "initial", "ansatz": ansatz, "noise": amp, "seed": seed})
    return F


def lc_params(theory, n=48, m=None, xsize=12.0, **kw):
    """Physical defaults of the reference (twist, coup_PotE ~ 4.06, coup_eps ~ 22.1, e3/e1 = 2)."""
    m = m or n
    h = xsize / (n - 1)
    base = dict(xlen=n, ylen=m, xsize=xsize, ysize=h * (m - 1), unit_magnetization=True)
    base.update(kw)
    return theory.params.default_params(**base)


def lc_field(theory, p_i, p_f, seed=0, amp=1e-2, ansatz="bloch"):
    """Skyrmion ansatz plus noise; Phi set to a smooth bump so the depolarisation terms act."""
    nf, X, Y = int(p_i[4]), int(p_i[0]), int(p_i[1])
    F = np.zeros((nf, X, Y))
    V = np.zeros_like(F)
    theory.initial_config.initialize(Field=F, Velocity=V, grid=grid_of(p_i, p_f), p_i_h=p_i, p_f_h=p_f,
                                     config={"mode": "initial", "ansatz": ansatz, "noise": amp, "seed": seed})
    g = grid_of(p_i, p_f)
    F[3] = 0.2 * np.exp(-((g[0] - g[0].mean()) ** 2 + (g[1] - g[1].mean()) ** 2) / 5.0)
    return F
