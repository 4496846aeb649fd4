"""Ferromagnetic superconductor initial conditions on host arrays (reference:
ferromagnetic_superconductor/initial_config.py:141-260, 481-554).

``"ground"``: m = (0, 0, M0), A = 0, psi = (u1, 0).
``"initial"`` (default): one skyrmion-vortex pair at the lattice centre.  The
magnetisation profile is pi exp(-(2r/R)^2) for r <= R = max(xsize, ysize)/10
(else 0), with angle theta = atan2 in [0, 2 pi) minus the rotation; Bloch,
Neel or anti texture of winding skyrmion_number, rescaled to |m| = M0.  The
superconducting profile is tanh(0.2 r) for r <= max(xsize, ysize) (else 1):
A = ainf fs (-sin th, cos th, 0)/max(r, 1e-12), psi = u1 fs e^{i N th} with N
taken from the config's ``vortex_number`` (default 1), as the reference's
launch does.  Unknown ansatz names fall back to Bloch.
Optional ``"noise"`` + ``"seed"``: seeded uniform perturbation of every field,
after which m is rescaled to M0 again.
"""
from __future__ import annotations

import math

import numpy as np


def _rescale_m(Field, M0):
    s = np.sqrt(Field[0] ** 2 + Field[1] ** 2 + Field[2] ** 2)
    nz = s != 0.0
    for a in range(3):
        Field[a] = np.where(nz, Field[a] * (M0 / np.where(nz, s, 1.0)), Field[a])


def initialize(*, Field, Velocity, grid, p_i_h, p_f_h, config=None):
    cfg = config or {}
    mode = str(cfg.get("mode", "initial")).lower()
    Velocity[...] = 0.0
    M0, u1 = float(p_f_h[17]), float(p_f_h[11])
    if mode in ("ground", "uniform", "vacuum", "gs"):
        Field[...] = 0.0
        Field[2] = M0
        Field[6] = u1
    else:
        ans = str(cfg.get("ansatz", "bloch")).lower()
        if ans not in ("bloch", "neel", "anti"):
            ans = "bloch"
        rot = float(cfg.get("skyrmion_rotation", 0.0))
        N = float(cfg.get("vortex_number", 1.0))
        skN, ainf = float(p_f_h[18]), float(p_f_h[13])
        X, Y = Field.shape[1], Field.shape[2]
        xc = (grid[0, 0, 0] + grid[0, X - 1, Y - 1]) / 2.0
        yc = (grid[1, 0, 0] + grid[1, X - 1, Y - 1]) / 2.0
        gx, gy = grid[0] - xc, grid[1] - yc
        r = np.sqrt(gx * gx + gy * gy)
        ang = np.arctan2(gy, gx)
        ang = np.where(ang < 0.0, ang + 2.0 * math.pi, ang)
        th = np.where(r == 0.0, -rot, ang - rot)
        xsize, ysize = float(p_f_h[0]), float(p_f_h[1])
        rmax = xsize if xsize > ysize else ysize
        R = rmax / 10.0
        t = 2.0 * r / R
        fm = np.where(r > R, 0.0, math.pi * np.exp(-(t * t)))
        fs = np.where(r > rmax, 1.0, np.tanh(0.2 * r))
        sf, cf = np.sin(fm), np.cos(fm)
        if ans == "bloch":
            Field[0], Field[1] = -sf * np.sin(skN * th), sf * np.cos(skN * th)
        elif ans == "neel":
            Field[0], Field[1] = sf * np.cos(skN * th), sf * np.sin(skN * th)
        else:
            Field[0], Field[1] = -sf * np.sin(skN * th), -sf * np.cos(skN * th)
        Field[2] = cf
        _rescale_m(Field, M0)
        rr = np.maximum(r, 1e-12)
        Field[3] = -ainf * fs * np.sin(th) / rr
        Field[4] = ainf * fs * np.cos(th) / rr
        Field[5] = 0.0
        Field[6] = u1 * fs * np.cos(N * th)
        Field[7] = u1 * fs * np.sin(N * th)
    amp = float(cfg.get("noise", 0.0))
    if amp > 0.0:
        rng = np.random.default_rng(int(cfg.get("seed", 0)))
        Field += rng.uniform(-amp, amp, size=Field.shape)
        _rescale_m(Field, M0)


# -- device generation ---------------------------------------------------
_MODE = {"ground": 0, "uniform": 0, "vacuum": 0, "gs": 0, "skyrmions": 2, "vortex": 3}
_ANSATZ = {"bloch": 0, "neel": 1, "anti": 2}


def device_plan(cfg, p_i_h, p_f_h):
    """The same configuration for ss_init_field / ss_add_noise (m rescaled to M0 after the noise)."""
    cfg = cfg or {}
    mode = _MODE.get(str(cfg.get("mode", "initial")).lower(), 1)
    desc = dict(mode=0 if mode == 0 else 1, ansatz=_ANSATZ.get(str(cfg.get("ansatz", "bloch")).lower(), 0),
                rotation=float(cfg.get("skyrmion_rotation", 0.0)), vortex_number=float(cfg.get("vortex_number", 1.0)))
    amp = float(cfg.get("noise", 0.0))
    noise = [(0, 8, 0, amp, 3, float(p_f_h[17]))] if amp > 0.0 else []
    return desc, noise
