"""Ginzburg-Landau superconductor / Abelian Higgs (oracle, FP64).

TEST INFRASTRUCTURE ONLY.  Restates soliton_solver/theories/
ginzburg_landau_superconductor/kernels.py.  Fields: A1, A2, psi1, psi2.
Packed parameters (ginzburg_landau_superconductor/params.py:161-194):
p_i[10] nhiggs, [11] ngauge; p_f[6] q, [7] Lambda, [8] u1,
[9] vortex_number, [10] ainf.
"""
from __future__ import annotations

import math

import numpy as np

NAME = "Ginzburg-Landau superconductor"
HAS_CONSTRAINT = False


def energy_density(F, D1, p_i, p_f):
    """Per-site energy times gv (kernels.py:45-108).

    Quirk kept: the energy reads ``u`` from p_f[9] (vortex_number) while the
    variation reads p_f[8] (u1) (kernels.py:76 vs :379).
    """
    q, lam, u1, gv = float(p_f[6]), float(p_f[7]), float(p_f[9]), float(p_f[4])
    A1, A2, psi1, psi2 = F[0], F[1], F[2], F[3]
    dpsi1dx, dpsi1dy = D1[2, 0], D1[2, 1]
    dpsi2dx, dpsi2dy = D1[3, 0], D1[3, 1]
    dA1dy = D1[0, 1]
    dA2dx = D1[1, 0]
    psi_sq = psi1 * psi1 + psi2 * psi2
    A_sq = A1 * A1 + A2 * A2
    e = 0.0 * F[0]
    e = e + 0.5 * (dpsi1dx * dpsi1dx + dpsi1dy * dpsi1dy + dpsi2dx * dpsi2dx + dpsi2dy * dpsi2dy)
    e = e + q * A1 * (psi1 * dpsi2dx - psi2 * dpsi1dx)
    e = e + q * A2 * (psi1 * dpsi2dy - psi2 * dpsi1dy)
    e = e + 0.5 * q * q * A_sq * psi_sq
    e = e + lam / 8.0 * (u1 * u1 - psi_sq) * (u1 * u1 - psi_sq)
    curlA = dA2dx - dA1dy
    e = e + 0.5 * curlA * curlA
    return e * gv


def variation(F, D1, D2, p_i, p_f):
    """Reference energy gradient (kernels.py:342-419); no mask, no projection."""
    q, lam, u1 = float(p_f[6]), float(p_f[7]), float(p_f[8])
    A1, A2, psi1, psi2 = F[0], F[1], F[2], F[3]
    psi_sq = psi1 * psi1 + psi2 * psi2
    dA1dx, dA1dy = D1[0, 0], D1[0, 1]
    dA2dx, dA2dy = D1[1, 0], D1[1, 1]
    dpsi1dx, dpsi1dy = D1[2, 0], D1[2, 1]
    dpsi2dx, dpsi2dy = D1[3, 0], D1[3, 1]
    d2A1d2y = D2[0, 1]
    d2A2d2x = D2[1, 0]
    d2psi1d2x, d2psi1d2y = D2[2, 0], D2[2, 1]
    d2psi2d2x, d2psi2d2y = D2[3, 0], D2[3, 1]
    d2A1dxdy = D2[0, 2]
    d2A2dxdy = D2[1, 2]
    EG = np.empty_like(F)
    EG[0] = (-d2A1d2y + d2A2dxdy + q * (psi1 * dpsi2dx - psi2 * dpsi1dx) + q * q * A1 * psi_sq)
    EG[1] = (-d2A2d2x + d2A1dxdy + q * (psi1 * dpsi2dy - psi2 * dpsi1dy) + q * q * A2 * psi_sq)
    EG[2] = (-d2psi1d2x - d2psi1d2y + 2.0 * q * (A1 * dpsi2dx + A2 * dpsi2dy) + q * psi2 * (dA1dx + dA2dy)
             + q * q * (A1 * A1 + A2 * A2) * psi1 - lam / 2.0 * (u1 * u1 - psi_sq) * psi1)
    EG[3] = (-d2psi2d2x - d2psi2d2y - 2.0 * q * (A1 * dpsi1dx + A2 * dpsi1dy) - q * psi1 * (dA1dx + dA2dy)
             + q * q * (A1 * A1 + A2 * A2) * psi2 - lam / 2.0 * (u1 * u1 - psi_sq) * psi2)
    return EG


def vortex_density(F, D1, p_i, p_f):
    """``(d_x A2 - d_y A1) * gv / (2 pi q)`` (kernels.py:145-178)."""
    gv, q = float(p_f[4]), float(p_f[6])
    b = D1[1, 0] - D1[0, 1]
    return b * (gv / (2.0 * math.pi * q))


CHARGES = {"vortex_number": vortex_density}
