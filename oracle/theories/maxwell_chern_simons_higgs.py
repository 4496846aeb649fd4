"""Maxwell-Chern-Simons-Higgs (anyon superconductor) (oracle, FP64).

TEST INFRASTRUCTURE ONLY.  Restates soliton_solver/theories/
maxwell_chern_simons_higgs/kernels.py.  Fields: A1, A2, psi1, psi2, A0 (= V).
Packed parameters (maxwell_chern_simons_higgs/params.py:173-207):
p_i[10] nhiggs, [11] ngauge; p_f[6] q, [7] Lambda, [8] kappa, [9] u1,
[10] vortex_number, [11] ainf.
"""
from __future__ import annotations

import math

import numpy as np

NAME = "Anyon superconductor"
HAS_CONSTRAINT = False


def energy_density(F, D1, p_i, p_f):
    """kernels.py:49-116 (A^2 includes A0^2; no Chern-Simons term)."""
    q, lam, u1, gv = float(p_f[6]), float(p_f[7]), float(p_f[9]), float(p_f[4])
    A1, A2, p1, p2, A0 = F[0], F[1], F[2], F[3], F[4]
    psq = p1 * p1 + p2 * p2
    Asq = A1 * A1 + A2 * A2 + A0 * A0
    e = 0.0 * F[0]
    e = e + 0.5 * (D1[2, 0] * D1[2, 0] + D1[2, 1] * D1[2, 1] + D1[3, 0] * D1[3, 0] + D1[3, 1] * D1[3, 1])
    e = e + q * A1 * (p1 * D1[3, 0] - p2 * D1[2, 0])
    e = e + q * A2 * (p1 * D1[3, 1] - p2 * D1[2, 1])
    e = e + 0.5 * q * q * Asq * psq
    e = e + lam / 8.0 * (u1 * u1 - psq) * (u1 * u1 - psq)
    curl = D1[1, 0] - D1[0, 1]
    e = e + 0.5 * curl * curl
    e = e + 0.5 * (D1[4, 0] * D1[4, 0] + D1[4, 1] * D1[4, 1])
    return e * gv


def variation(F, D1, D2, p_i, p_f):
    """kernels.py:480-567 (saddle field A0: -V^2 in the psi equations)."""
    q, lam, kappa, u1 = float(p_f[6]), float(p_f[7]), float(p_f[8]), float(p_f[9])
    A1, A2, p1, p2, V = F[0], F[1], F[2], F[3], F[4]
    psq = p1 * p1 + p2 * p2
    dA1dx, dA1dy = D1[0, 0], D1[0, 1]
    dA2dx, dA2dy = D1[1, 0], D1[1, 1]
    dp1dx, dp1dy = D1[2, 0], D1[2, 1]
    dp2dx, dp2dy = D1[3, 0], D1[3, 1]
    dVdx, dVdy = D1[4, 0], D1[4, 1]
    g0 = (-D2[0, 1] + D2[1, 2] + q * (p1 * dp2dx - p2 * dp1dx) + q * q * A1 * psq - kappa * dVdy)
    g1 = (-D2[1, 0] + D2[0, 2] + q * (p1 * dp2dy - p2 * dp1dy) + q * q * A2 * psq + kappa * dVdx)
    g2 = (-D2[2, 0] - D2[2, 1] + 2.0 * q * (A1 * dp2dx + A2 * dp2dy) + q * p2 * (dA1dx + dA2dy)
          + q * q * (A1 * A1 + A2 * A2 - V * V) * p1 - lam / 2.0 * (u1 * u1 - psq) * p1)
    g3 = (-D2[3, 0] - D2[3, 1] - 2.0 * q * (A1 * dp1dx + A2 * dp1dy) - q * p1 * (dA1dx + dA2dy)
          + q * q * (A1 * A1 + A2 * A2 - V * V) * p2 - lam / 2.0 * (u1 * u1 - psq) * p2)
    g4 = (-D2[4, 0] - D2[4, 1] + q * q * V * psq + kappa * (dA2dx - dA1dy))
    return np.stack([g0, g1, g2, g3, g4])


def vortex_density(F, D1, p_i, p_f):
    """kernels.py:153-186."""
    gv, q = float(p_f[4]), float(p_f[6])
    return (D1[1, 0] - D1[0, 1]) * (gv / (2.0 * math.pi * q))


def electric_charge_density(F, D1, D2, p_i, p_f):
    """-(d2 A0/dx2 + d2 A0/dy2) gv (kernels.py:350-380)."""
    return -(D2[4, 0] + D2[4, 1]) * float(p_f[4])


def noether_charge_density(F, D1, p_i, p_f):
    """q^2 A0 |psi|^2 gv (kernels.py:417-478)."""
    q, gv = float(p_f[6]), float(p_f[4])
    return q * q * F[4] * (F[2] * F[2] + F[3] * F[3]) * gv


CHARGES = {"vortex_number": vortex_density, "noether_charge": noether_charge_density}
