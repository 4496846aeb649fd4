"""Spin-triplet superconducting ferromagnet initial conditions on host arrays
(reference: spin_triplet_superconducting_magnet/initial_config.py, ground-state
and initial-configuration kernels and ``initialize``).

``"ground"``: m = (0, 0, M0), A = 0, psi1 = (u1, 0), psi2 = (u2, 0).
``"initial"`` (default): one skyrmion and a two-component vortex at the lattice
centre: magnetisation as in the ferromagnetic superconductor (profile
pi exp(-(2r/R)^2), R = max(xsize, ysize)/10, Bloch/Neel/anti texture of winding
skyrmion_number, rescaled to |m| = M0), A = ainf fs (-sin th, cos th, 0)/r and
psi_k = u_k fs e^{i N_k th} with N_k = vortex1_number, vortex2_number from the
parameters and fs = tanh(0.2 r).  Optional ``"noise"`` + ``"seed"``: seeded
uniform perturbation of every field, then |m| rescaled to M0.
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
    M0, u1, u2 = float(p_f_h[20]), float(p_f_h[11]), float(p_f_h[12])
    if mode in ("ground", "uniform", "vacuum", "gs"):
        Field[...] = 0.0
        Field[2] = M0
        Field[6] = u1
        Field[8] = u2
    else:
        ans = str(cfg.get("ansatz", "bloch")).lower()
        if ans not in ("bloch", "neel", "anti"):
            ans = "bloch"
        rot = float(cfg.get("skyrmion_rotation", 0.0))
        skN, ainf = float(p_f_h[21]), float(p_f_h[16])
        N1, N2 = float(p_f_h[13]), float(p_f_h[14])
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
        Field[6] = u1 * fs * np.cos(N1 * th)
        Field[7] = u1 * fs * np.sin(N1 * th)
        Field[8] = u2 * fs * np.cos(N2 * th)
        Field[9] = u2 * fs * np.sin(N2 * th)
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
                rotation=float(cfg.get("skyrmion_rotation", 0.0)))
    amp = float(cfg.get("noise", 0.0))
    noise = [(0, 10, 0, amp, 3, float(p_f_h[20]))] if amp > 0.0 else []
    return desc, noise
