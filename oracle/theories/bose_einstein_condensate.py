"""Rotating Bose-Einstein condensate (oracle, FP64).

TEST INFRASTRUCTURE ONLY.  Restates soliton_solver/theories/
bose_einstein_condensate/kernels.py and observables.py.  Fields: psi_r, psi_i.
Packed parameters (bose_einstein_condensate/params.py:198-229): p_i[10] nhiggs;
p_f[6] beta, [7] tf_radius, [8] omega_rot, [9] vortex_number.
The trap centre is index based: ((lsx (xlen-1))/2, (lsy (ylen-1))/2)
(kernels.py:74-80).
"""
from __future__ import annotations

import math

import numpy as np

from oracle import core

NAME = "Bose-Einstein condensate"
HAS_CONSTRAINT = False


def _positions(p_i, p_f):
    X, Y = int(p_i[0]), int(p_i[1])
    lsx, lsy = float(p_f[2]), float(p_f[3])
    xcent = (0.0 + lsx * float(X - 1)) / 2.0
    ycent = (0.0 + lsy * float(Y - 1)) / 2.0
    xpos = lsx * np.arange(X, dtype=np.float64)[:, None] - xcent
    ypos = lsy * np.arange(Y, dtype=np.float64)[None, :] - ycent
    xpos, ypos = np.broadcast_arrays(xpos, ypos)
    return xpos, ypos, np.sqrt(xpos * xpos + ypos * ypos)


def energy_density(F, D1, p_i, p_f):
    """kernels.py:45-102."""
    beta, omega, gv = float(p_f[6]), float(p_f[8]), float(p_f[4])
    xpos, ypos, r1 = _positions(p_i, p_f)
    pr, pi = F[0], F[1]
    dprdx, dprdy = D1[0, 0], D1[0, 1]
    dpidx, dpidy = D1[1, 0], D1[1, 1]
    psq = pr * pr + pi * pi
    e = 0.0 * F[0]
    e = e + 0.5 * (dprdx * dprdx + dprdy * dprdy + dpidx * dpidx + dpidy * dpidy)
    e = e + 0.5 * r1 * r1 * psq
    e = e + 0.5 * beta * psq * psq
    e = e - omega * (xpos * (pr * dpidy - pi * dprdy) - ypos * (pr * dpidx - pi * dprdx))
    return e * gv


def variation(F, D1, D2, p_i, p_f):
    """kernels.py:256-328; no mask, no projection."""
    beta, omega = float(p_f[6]), float(p_f[8])
    xpos, ypos, r1 = _positions(p_i, p_f)
    pr, pi = F[0], F[1]
    dprdx, dprdy = D1[0, 0], D1[0, 1]
    dpidx, dpidy = D1[1, 0], D1[1, 1]
    psq = pr * pr + pi * pi
    g0 = pr * r1 * r1 + 2.0 * beta * pr * psq - D2[0, 0] - D2[0, 1]
    g1 = pi * r1 * r1 + 2.0 * beta * pi * psq - D2[1, 0] - D2[1, 1]
    g0 = g0 - 2.0 * omega * (xpos * dpidy - ypos * dpidx)
    g1 = g1 + 2.0 * omega * (xpos * dprdy - ypos * dprdx)
    return np.stack([g0, g1])


def l2_normalize(G, p_i, p_f):
    """compute_norm (observables.py:58-96: sum |psi|^2 gv in the reference
    order) then psi /= sqrt(norm) (kernels.py:196-254), in place."""
    gv = float(p_f[4])
    norm = core.reference_sum(((G[0] * G[0] + G[1] * G[1]) * gv).ravel())
    s = math.sqrt(norm)
    G[0] = G[0] / s
    G[1] = G[1] / s


def norm_density(F, D1, p_i, p_f):
    return (F[0] * F[0] + F[1] * F[1]) * float(p_f[4])


CHARGES = {}
