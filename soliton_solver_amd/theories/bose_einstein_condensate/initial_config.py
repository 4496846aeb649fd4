"""Bose-Einstein condensate initial conditions on host arrays (reference:
bose_einstein_condensate/initial_config.py:143-515).

``"ground"`` / ``"initial"``: Thomas-Fermi amplitude sqrt(max(0, 1/sqrt(beta pi)
- r^2/(2 beta))) for r <= tf_radius about the lattice centre, times
|vortex_number| vortices seeded on hexagonal rings of spacing tf_radius/5
(phase sign(N) theta_n, core tanh(r_n/(sqrt(2) xi)), xi = spacing/2); the
vortex count is p_f[9] for "ground" and ``config["vortex_number"]`` otherwise.
``"noise"`` + ``"seed"`` add a seeded uniform perturbation.
"""
from __future__ import annotations

import math

import numpy as np


def _seed_position(n, xc, yc, spacing):
    if n == 0:
        return xc, yc
    remaining, ring = n - 1, 1
    while True:
        cap = 6 * ring
        if remaining < cap:
            th = 2.0 * math.pi * (remaining / cap)
            return xc + ring * spacing * math.cos(th), yc + ring * spacing * math.sin(th)
        remaining -= cap
        ring += 1


def initialize(*, Field, Velocity, grid, p_i_h, p_f_h, config=None):
    cfg = config or {}
    mode = str(cfg.get("mode", "initial")).lower()
    Velocity[...] = 0.0
    beta, tf_radius = float(p_f_h[6]), float(p_f_h[7])
    nv_signed = float(p_f_h[9]) if mode in ("ground", "uniform", "vacuum", "gs") else float(cfg.get("vortex_number", 1.0))
    nv = int(abs(nv_signed))
    sign = 1.0 if nv_signed > 0 else (-1.0 if nv_signed < 0 else 0.0)
    X, Y = Field.shape[1], Field.shape[2]
    xc = (grid[0, 0, 0] + grid[0, X - 1, Y - 1]) / 2.0
    yc = (grid[1, 0, 0] + grid[1, X - 1, Y - 1]) / 2.0
    dx0, dy0 = grid[0] - xc, grid[1] - yc
    rtf = np.sqrt(dx0 * dx0 + dy0 * dy0)
    tf = np.where(rtf <= tf_radius, 1.0 / math.sqrt(beta * math.pi) - 1.0 / (2.0 * beta) * rtf * rtf, 0.0)
    amp = np.sqrt(np.maximum(tf, 0.0))
    spacing = tf_radius / 5.0
    xi = spacing / 2.0
    phase = np.zeros_like(amp)
    core = np.ones_like(amp)
    for n in range(nv):
        xv, yv = _seed_position(n, xc, yc, spacing)
        dx, dy = grid[0] - xv, grid[1] - yv
        phase = phase + sign * np.arctan2(dy, dx)
        core = core * np.tanh(np.sqrt(dx * dx + dy * dy) / (math.sqrt(2.0) * xi))
    Field[0] = amp * core * np.cos(phase)
    Field[1] = amp * core * np.sin(phase)
    a = float(cfg.get("noise", 0.0))
    if a > 0.0:
        rng = np.random.default_rng(int(cfg.get("seed", 0)))
        Field += rng.uniform(-a, a, size=Field.shape)


# -- device generation ---------------------------------------------------
_MODE = {"ground": 0, "uniform": 0, "vacuum": 0, "gs": 0, "skyrmions": 2, "vortex": 3}
_ANSATZ = {"bloch": 0, "neel": 1, "anti": 2}


def device_plan(cfg, p_i_h, p_f_h):
    """The same configuration for ss_init_field / ss_add_noise: Thomas-Fermi radius,
    beta, vortex sign and core size in ``extra``, vortex seeds as the soliton list."""
    cfg = cfg or {}
    mode = str(cfg.get("mode", "initial")).lower()
    beta, tf_radius = float(p_f_h[6]), float(p_f_h[7])
    nv_signed = float(p_f_h[9]) if mode in ("ground", "uniform", "vacuum", "gs") else float(cfg.get("vortex_number", 1.0))
    nv = int(abs(nv_signed))
    sign = 1.0 if nv_signed > 0 else (-1.0 if nv_signed < 0 else 0.0)
    X, Y = int(p_i_h[0]), int(p_i_h[1])
    xc, yc = (0.0 + float(p_f_h[2]) * (X - 1)) / 2.0, (0.0 + float(p_f_h[3]) * (Y - 1)) / 2.0
    spacing = tf_radius / 5.0
    seeds = [(*_seed_position(n, xc, yc, spacing), 0.0) for n in range(nv)]
    desc = dict(mode=1, extra=(tf_radius, beta, sign, spacing / 2.0), solitons=seeds)
    amp = float(cfg.get("noise", 0.0))
    noise = [(0, 2, 0, amp, 0, 0.0)] if amp > 0.0 else []
    return desc, noise
