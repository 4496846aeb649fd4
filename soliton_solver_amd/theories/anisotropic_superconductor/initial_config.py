"""Anisotropic superconductor initial conditions on host arrays (reference:
anisotropic_superconductor/initial_config.py:117-280).

``"ground"``: A = 0, psi1 = u1, psi2 = u2 e^{i theta_diff}.  The reference's
ground kernel reads u1, u2 from p_f[11], p_f[12] (beta2, beta3 in its own
packing); here the packed vacuum values p_f[17], p_f[18] are used.
``"initial"``: centred vortex, profile tanh(r) for r <= max(xsize, ysize),
A = ainf f (sin th, -cos th)/r, psi_k = u_k f e^{i N_k th}.
Optional ``"noise"`` + ``"seed"``: seeded uniform perturbation of all fields.
"""
from __future__ import annotations

import math

import numpy as np


def initialize(*, Field, Velocity, grid, p_i_h, p_f_h, config=None):
    cfg = config or {}
    mode = str(cfg.get("mode", "initial")).lower()
    Velocity[...] = 0.0
    u1, u2, th_d = float(p_f_h[17]), float(p_f_h[18]), float(p_f_h[19])
    if mode in ("ground", "uniform", "vacuum", "gs"):
        Field[0] = 0.0
        Field[1] = 0.0
        Field[2] = u1
        Field[3] = 0.0
        Field[4] = u2 * math.cos(th_d)
        Field[5] = u2 * math.sin(th_d)
    else:
        X, Y = Field.shape[1], Field.shape[2]
        xc = (grid[0, 0, 0] + grid[0, X - 1, Y - 1]) / 2.0
        yc = (grid[1, 0, 0] + grid[1, X - 1, Y - 1]) / 2.0
        gx, gy = grid[0] - xc, grid[1] - yc
        r = np.sqrt(gx * gx + gy * gy)
        th = np.arctan2(gy, gx)
        th = np.where(th < 0.0, th + 2.0 * np.pi, th)
        th = np.where(r == 0.0, 0.0, th)
        xsize, ysize = float(p_f_h[0]), float(p_f_h[1])
        rmax = xsize if xsize > ysize else ysize
        f = np.where(r > rmax, 1.0, np.tanh(r))
        rr = np.maximum(r, 1e-12)
        ainf, N1, N2 = float(p_f_h[22]), float(p_f_h[23]), float(p_f_h[24])
        Field[0] = ainf * f * np.sin(th) / rr
        Field[1] = -ainf * f * np.cos(th) / rr
        Field[2] = u1 * f * np.cos(N1 * th)
        Field[3] = u1 * f * np.sin(N1 * th)
        Field[4] = u2 * f * np.cos(N2 * th)
        Field[5] = u2 * f * np.sin(N2 * th)
    amp = float(cfg.get("noise", 0.0))
    if amp > 0.0:
        rng = np.random.default_rng(int(cfg.get("seed", 0)))
        Field += rng.uniform(-amp, amp, size=Field.shape)


# -- device generation ---------------------------------------------------
_MODE = {"ground": 0, "uniform": 0, "vacuum": 0, "gs": 0, "skyrmions": 2, "vortex": 3}
_ANSATZ = {"bloch": 0, "neel": 1, "anti": 2}


def device_plan(cfg, p_i_h, p_f_h):
    """The same configuration for ss_init_field / ss_add_noise."""
    cfg = cfg or {}
    mode = _MODE.get(str(cfg.get("mode", "initial")).lower(), 1)
    desc = dict(mode=0 if mode == 0 else 1)
    amp = float(cfg.get("noise", 0.0))
    noise = [(0, 6, 0, amp, 0, 0.0)] if amp > 0.0 else []
    return desc, noise
