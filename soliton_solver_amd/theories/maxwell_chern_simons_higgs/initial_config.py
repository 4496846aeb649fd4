"""Maxwell-Chern-Simons-Higgs initial conditions on host arrays (reference:
maxwell_chern_simons_higgs/initial_config.py): ``"ground"`` (A = 0, psi = (u1, 0),
A0 = 0) and ``"initial"`` (centred vortex with profile tanh(0.2 r) for
r <= max(xsize, ysize), A = ainf f (sin th, -cos th)/r, psi = u1 f e^{i N th},
A0 = 0).  ``"noise"`` + ``"seed"`` add a seeded uniform perturbation to all fields."""
from __future__ import annotations

import numpy as np


def initialize(*, Field, Velocity, grid, p_i_h, p_f_h, config=None):
    cfg = config or {}
    mode = str(cfg.get("mode", "initial")).lower()
    Velocity[...] = 0.0
    u1, ainf = float(p_f_h[9]), float(p_f_h[11])
    Field[...] = 0.0
    if mode in ("ground", "uniform", "vacuum", "gs"):
        Field[2] = u1
    else:
        X, Y = Field.shape[1], Field.shape[2]
        xc = (grid[0, 0, 0] + grid[0, X - 1, Y - 1]) / 2.0
        yc = (grid[1, 0, 0] + grid[1, X - 1, Y - 1]) / 2.0
        gx, gy = grid[0] - xc, grid[1] - yc
        r = np.sqrt(gx * gx + gy * gy)
        th = np.arctan2(gy, gx)
        th = np.where(th < 0.0, th + 2.0 * np.pi, th)
        th = np.where(r == 0.0, 0.0, th)
        rmax = max(float(p_f_h[0]), float(p_f_h[1]))
        f = np.where(r > rmax, 1.0, np.tanh(0.2 * r))
        N = float(cfg.get("vortex_number", 1.0))
        rr = np.maximum(r, 1e-12)
        Field[0] = ainf * f * np.sin(th) / rr
        Field[1] = -ainf * f * np.cos(th) / rr
        Field[2] = u1 * f * np.cos(N * th)
        Field[3] = u1 * f * np.sin(N * th)
    a = float(cfg.get("noise", 0.0))
    if a > 0.0:
        rng = np.random.default_rng(int(cfg.get("seed", 0)))
        Field += rng.uniform(-a, a, size=Field.shape)


# -- device generation ---------------------------------------------------
_MODE = {"ground": 0, "uniform": 0, "vacuum": 0, "gs": 0, "skyrmions": 2, "vortex": 3}
_ANSATZ = {"bloch": 0, "neel": 1, "anti": 2}


def device_plan(cfg, p_i_h, p_f_h):
    """The same configuration for ss_init_field / ss_add_noise."""
    cfg = cfg or {}
    mode = _MODE.get(str(cfg.get("mode", "initial")).lower(), 1)
    desc = dict(mode=0 if mode == 0 else 1, vortex_number=float(cfg.get("vortex_number", 1.0)))
    amp = float(cfg.get("noise", 0.0))
    noise = [(0, 5, 0, amp, 0, 0.0)] if amp > 0.0 else []
    return desc, noise
