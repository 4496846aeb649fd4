"""Chiral magnet initial conditions, on host arrays (reference:
theories/chiral_magnet/initial_config.py:117-460).

Modes: ``"ground"`` (m = e3, psi = 0; :117-180), ``"initial"`` (one ansatz at
the lattice centre with radius scale max(xsize, ysize)/10 * skyrmion_number;
:182-283 -- the reference's launch of this kernel passes one argument too few,
here the intended formula is evaluated), ``"skyrmions"`` (ground state plus a
list of placed solitons, :285-386).  Optional ``"noise"`` (sigma) with
``"seed"`` adds a seeded Gaussian perturbation to m and renormalises.
"""
from __future__ import annotations

import numpy as np

from ...core import ansatz as A


def _ground(Field):
    Field[0] = 0.0
    Field[1] = 0.0
    Field[2] = 1.0
    Field[3] = 0.0


def initialize(*, Field, Velocity, grid, p_i_h, p_f_h, config=None):
    cfg = config or {}
    mode = str(cfg.get("mode", "initial")).lower()
    Velocity[...] = 0.0
    xsize, ysize = float(p_f_h[0]), float(p_f_h[1])
    rmax = xsize if xsize > ysize else ysize
    skN = float(p_f_h[9])
    kind = str(cfg.get("ansatz", "bloch")).lower()
    if kind not in ("bloch", "neel", "anti"):
        kind = "bloch"
    rot = float(cfg.get("skyrmion_rotation", 0.0))
    if mode in ("ground", "uniform", "vacuum", "gs"):
        _ground(Field)
    elif mode == "skyrmions":
        _ground(Field)
        for sk in cfg.get("skyrmions", ()):
            xc, yc, radius = float(sk[0]), float(sk[1]), float(sk[2])
            A.place_soliton(Field, grid, xc, yc, 2.0 * radius, rot, skN, kind)
    else:
        X, Y = Field.shape[1], Field.shape[2]
        xcent = (grid[0, 0, 0] + grid[0, X - 1, Y - 1]) / 2.0
        ycent = (grid[1, 0, 0] + grid[1, X - 1, Y - 1]) / 2.0
        A.single_ansatz(Field, grid, xcent, ycent, rmax / 10.0 * skN, rot, skN, kind)
        Field[3] = 0.0
    sigma = float(cfg.get("noise", 0.0))
    if sigma > 0.0:
        A.gaussian_perturb(Field, np.random.default_rng(int(cfg.get("seed", 0))), sigma, 3)


# -- device generation ---------------------------------------------------
_MODE = {"ground": 0, "uniform": 0, "vacuum": 0, "gs": 0, "skyrmions": 2, "vortex": 3}
_ANSATZ = {"bloch": 0, "neel": 1, "anti": 2}


def device_plan(cfg, p_i_h, p_f_h):
    """The same configuration for ss_init_field / ss_add_noise (noise: Gaussian on m, renormalised)."""
    cfg = cfg or {}
    mode = _MODE.get(str(cfg.get("mode", "initial")).lower(), 1)
    mode = mode if mode in (0, 2) else 1
    sol = [(float(s[0]), float(s[1]), float(s[2])) for s in cfg.get("skyrmions", ())] if mode == 2 else []
    desc = dict(mode=mode, ansatz=_ANSATZ.get(str(cfg.get("ansatz", "bloch")).lower(), 0),
                skyrmion_number=float(p_f_h[9]), rotation=float(cfg.get("skyrmion_rotation", 0.0)), solitons=sol)
    sigma = float(cfg.get("noise", 0.0))
    noise = [(0, 3, 1, sigma, 3, 1.0)] if sigma > 0.0 else []
    return desc, noise
