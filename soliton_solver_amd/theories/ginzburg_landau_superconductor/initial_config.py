"""Ginzburg-Landau initial conditions on host arrays (reference:
ginzburg_landau_superconductor/initial_config.py:114-379).

``"ground"``: A = 0, psi = (u1, 0) (:114-178).
``"initial"``: centred vortex with profile tanh(0.2 r) for r <= max(xsize, ysize),
A = ainf * f (sin th, -cos th)/r and psi = u1 f e^{i N th} (:180-257); the
reference reads ainf from p_f[11], one past the end of its 11-entry packing,
so here ainf is p_f[10] as packed.
``"vortex"``: psi = u1 tanh(scale * r) e^{i N th} about the centre with A = 0,
plus optional seeded uniform noise U(-noise, noise) on Re and Im psi.
"""
from __future__ import annotations

import numpy as np


def _centre(grid):
    X, Y = grid.shape[1], grid.shape[2]
    return (grid[0, 0, 0] + grid[0, X - 1, Y - 1]) / 2.0, (grid[1, 0, 0] + grid[1, X - 1, Y - 1]) / 2.0


def _polar(grid, xc, yc):
    gx, gy = grid[0] - xc, grid[1] - yc
    r = np.sqrt(gx * gx + gy * gy)
    ang = np.arctan2(gy, gx)
    ang = np.where(ang < 0.0, ang + 2.0 * np.pi, ang)
    return r, np.where(r == 0.0, 0.0, ang)


def initialize(*, Field, Velocity, grid, p_i_h, p_f_h, config=None):
    cfg = config or {}
    mode = str(cfg.get("mode", "initial")).lower()
    Velocity[...] = 0.0
    u1 = float(p_f_h[8])
    if mode in ("ground", "uniform", "vacuum", "gs"):
        Field[0] = 0.0
        Field[1] = 0.0
        Field[2] = u1
        Field[3] = 0.0
        return
    N = float(cfg.get("vortex_number", 1.0))
    xc, yc = _centre(grid)
    r, th = _polar(grid, xc, yc)
    if mode == "vortex":
        f = np.tanh(float(cfg.get("scale", 1.0)) * r)
        Field[0] = 0.0
        Field[1] = 0.0
        Field[2] = u1 * f * np.cos(N * th)
        Field[3] = u1 * f * np.sin(N * th)
        amp = float(cfg.get("noise", 0.0))
        if amp > 0.0:
            rng = np.random.default_rng(int(cfg.get("seed", 0)))
            Field[2] += rng.uniform(-amp, amp, size=Field[2].shape)
            Field[3] += rng.uniform(-amp, amp, size=Field[3].shape)
        return
    xsize, ysize = float(p_f_h[0]), float(p_f_h[1])
    rmax = xsize if xsize > ysize else ysize
    ainf = float(p_f_h[10])
    f = np.where(r > rmax, 1.0, np.tanh(0.2 * r))
    rr = np.maximum(r, 1e-12)
    Field[0] = ainf * f * np.sin(th) / rr
    Field[1] = -ainf * f * np.cos(th) / rr
    Field[2] = u1 * f * np.cos(N * th)
    Field[3] = u1 * f * np.sin(N * th)


# -- device generation ---------------------------------------------------
_MODE = {"ground": 0, "uniform": 0, "vacuum": 0, "gs": 0, "skyrmions": 2, "vortex": 3}
_ANSATZ = {"bloch": 0, "neel": 1, "anti": 2}


def device_plan(cfg, p_i_h, p_f_h):
    """The same configuration for ss_init_field / ss_add_noise (noise on psi in the vortex mode only)."""
    cfg = cfg or {}
    mode = _MODE.get(str(cfg.get("mode", "initial")).lower(), 1)
    desc = dict(mode=mode if mode in (0, 3) else 1, vortex_number=float(cfg.get("vortex_number", 1.0)),
                scale=float(cfg.get("scale", 1.0)))
    amp = float(cfg.get("noise", 0.0))
    noise = [(2, 4, 0, amp, 0, 0.0)] if (amp > 0.0 and mode == 3) else []
    return desc, noise
