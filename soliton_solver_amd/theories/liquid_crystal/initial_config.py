"""Liquid crystal initial conditions on host arrays (reference:
liquid_crystal/initial_config.py, ground-state and initial-configuration kernels
and ``initialize``).

``"ground"``: n = e3, Phi = 0.  ``"initial"`` (default): one skyrmion at the
lattice centre with profile pi exp(-(2r/R)^2) for r <= R = max(xsize, ysize)/10
times skyrmion_number (else 0), Bloch/Neel/anti texture, Phi = 0.  The
reference's launch of this kernel passes one argument too few; here the
intended formula is evaluated.  Optional ``"noise"`` + ``"seed"``: seeded
uniform perturbation of n (renormalised) and Phi.
"""
from __future__ import annotations

import math

import numpy as np


def _unit(Field):
    s = np.sqrt(Field[0] ** 2 + Field[1] ** 2 + Field[2] ** 2)
    nz = s != 0.0
    for a in range(3):
        Field[a] = np.where(nz, Field[a] / np.where(nz, s, 1.0), Field[a])


def initialize(*, Field, Velocity, grid, p_i_h, p_f_h, config=None):
    cfg = config or {}
    mode = str(cfg.get("mode", "initial")).lower()
    Velocity[...] = 0.0
    Field[...] = 0.0
    if mode in ("ground", "uniform", "vacuum", "gs"):
        Field[2] = 1.0
    else:
        ans = str(cfg.get("ansatz", "bloch")).lower()
        if ans not in ("bloch", "neel", "anti"):
            ans = "bloch"
        rot = float(cfg.get("skyrmion_rotation", 0.0))
        skN = float(p_f_h[11])
        X, Y = Field.shape[1], Field.shape[2]
        xc = (grid[0, 0, 0] + grid[0, X - 1, Y - 1]) / 2.0
        yc = (grid[1, 0, 0] + grid[1, X - 1, Y - 1]) / 2.0
        gx, gy = grid[0] - xc, grid[1] - yc
        r = np.sqrt(gx * gx + gy * gy)
        ang = np.arctan2(gy, gx)
        ang = np.where(ang < 0.0, ang + 2.0 * math.pi, ang)
        th = np.where(r == 0.0, -rot, ang - rot)
        xsize, ysize = float(p_f_h[0]), float(p_f_h[1])
        R = (xsize if xsize > ysize else ysize) / 10.0 * skN
        t = 2.0 * r / R
        fm = np.where(r > R, 0.0, math.pi * np.exp(-(t * t)))
        sf, cf = np.sin(fm), np.cos(fm)
        if ans == "bloch":
            Field[0], Field[1] = -sf * np.sin(skN * th), sf * np.cos(skN * th)
        elif ans == "neel":
            Field[0], Field[1] = sf * np.cos(skN * th), sf * np.sin(skN * th)
        else:
            Field[0], Field[1] = -sf * np.sin(skN * th), -sf * np.cos(skN * th)
        Field[2] = cf
        _unit(Field)
    amp = float(cfg.get("noise", 0.0))
    if amp > 0.0:
        rng = np.random.default_rng(int(cfg.get("seed", 0)))
        Field += rng.uniform(-amp, amp, size=Field.shape)
        _unit(Field)


# -- device generation ---------------------------------------------------
_MODE = {"ground": 0, "uniform": 0, "vacuum": 0, "gs": 0, "skyrmions": 2, "vortex": 3}
_ANSATZ = {"bloch": 0, "neel": 1, "anti": 2}


def device_plan(cfg, p_i_h, p_f_h):
    """The same configuration for ss_init_field / ss_add_noise."""
    cfg = cfg or {}
    mode = _MODE.get(str(cfg.get("mode", "initial")).lower(), 1)
    desc = dict(mode=0 if mode == 0 else 1, ansatz=_ANSATZ.get(str(cfg.get("ansatz", "bloch")).lower(), 0),
                skyrmion_number=float(p_f_h[11]), rotation=float(cfg.get("skyrmion_rotation", 0.0)))
    amp = float(cfg.get("noise", 0.0))
    noise = [(0, 4, 0, amp, 3, 1.0)] if amp > 0.0 else []
    return desc, noise
