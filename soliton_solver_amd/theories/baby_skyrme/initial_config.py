"""Baby Skyrme initial conditions on host arrays (reference:
theories/baby_skyrme/initial_config.py:120-444): ``"ground"`` (m = e3),
``"initial"`` (one Bloch / Neel / anti ansatz at the centre with radius scale
max(xsize, ysize)/10 * skyrmion_number, p_f[8]; the reference's launch of this
kernel passes one argument too few, here the intended formula is evaluated),
``"skyrmions"`` (placed solitons).  ``"noise"`` + ``"seed"``: seeded uniform
perturbation U(-noise, noise) of m, renormalised (``"noise_kind": "gauss"`` for
Gaussian)."""
from __future__ import annotations

import numpy as np

from ...core import ansatz as A


def initialize(*, Field, Velocity, grid, p_i_h, p_f_h, config=None):
    cfg = config or {}
    mode = str(cfg.get("mode", "initial")).lower()
    Velocity[...] = 0.0
    xsize, ysize = float(p_f_h[0]), float(p_f_h[1])
    rmax = xsize if xsize > ysize else ysize
    skN = float(cfg.get("skyrmion_number", p_f_h[8]))
    kind = str(cfg.get("ansatz", "bloch")).lower()
    if kind not in ("bloch", "neel", "anti"):
        kind = "bloch"
    rot = float(cfg.get("skyrmion_rotation", 0.0))
    Field[0] = 0.0
    Field[1] = 0.0
    Field[2] = 1.0
    if mode in ("ground", "uniform", "vacuum", "gs"):
        pass
    elif mode == "skyrmions":
        for sk in cfg.get("skyrmions", ()):
            A.place_soliton(Field, grid, float(sk[0]), float(sk[1]), 2.0 * float(sk[2]), rot, skN, kind)
    else:
        X, Y = Field.shape[1], Field.shape[2]
        xcent = (grid[0, 0, 0] + grid[0, X - 1, Y - 1]) / 2.0
        ycent = (grid[1, 0, 0] + grid[1, X - 1, Y - 1]) / 2.0
        A.single_ansatz(Field, grid, xcent, ycent, rmax / 10.0 * skN, rot, skN, kind)
    amp = float(cfg.get("noise", 0.0))
    if amp > 0.0:
        rng = np.random.default_rng(int(cfg.get("seed", 0)))
        if str(cfg.get("noise_kind", "uniform")) == "gauss":
            A.gaussian_perturb(Field, rng, amp, 3)
        else:
            for a in range(3):
                Field[a] = Field[a] + rng.uniform(-amp, amp, size=Field[a].shape)
            A.unit_normalize(Field, 3)


# -- device generation ---------------------------------------------------
_MODE = {"ground": 0, "uniform": 0, "vacuum": 0, "gs": 0, "skyrmions": 2, "vortex": 3}
_ANSATZ = {"bloch": 0, "neel": 1, "anti": 2}


def device_plan(cfg, p_i_h, p_f_h):
    """The same configuration for ss_init_field / ss_add_noise."""
    cfg = cfg or {}
    mode = _MODE.get(str(cfg.get("mode", "initial")).lower(), 1)
    mode = mode if mode in (0, 2) else 1
    sol = [(float(s[0]), float(s[1]), float(s[2])) for s in cfg.get("skyrmions", ())] if mode == 2 else []
    desc = dict(mode=mode, ansatz=_ANSATZ.get(str(cfg.get("ansatz", "bloch")).lower(), 0),
                skyrmion_number=float(cfg.get("skyrmion_number", p_f_h[8])),
                rotation=float(cfg.get("skyrmion_rotation", 0.0)), solitons=sol)
    amp = float(cfg.get("noise", 0.0))
    gauss = 1 if str(cfg.get("noise_kind", "uniform")) == "gauss" else 0
    noise = [(0, 3, gauss, amp, 3, 1.0)] if amp > 0.0 else []
    return desc, noise
