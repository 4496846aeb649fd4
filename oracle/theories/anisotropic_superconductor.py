"""Anisotropic s+id superconductor (oracle, FP64).

TEST INFRASTRUCTURE ONLY.  Restates soliton_solver/theories/
anisotropic_superconductor/kernels.py.  Fields: A1, A2, psi1 (re, im),
psi2 (re, im).  Packed parameters (anisotropic_superconductor/params.py:266-313):
p_f[6] kappa, [7] Q, [8] alpha1, [9] alpha2, [10..13] beta1..beta4,
[14] gamma1, [15] gamma2, [16] gamma12, [17] u1, [18] u2, [19] theta_diff,
[20] beta3tilde, [21] vortex_number, [22] ainf, [23] N1, [24] N2.
"""
from __future__ import annotations

import math

import numpy as np

NAME = "Anisotropic superconductor"
HAS_CONSTRAINT = False


def gamma_entry(i, j, a, b, p_f):
    """kernels.py:45-94: 2 gamma1 / -+2 gamma12 / 2 gamma2 on i == j, 0 otherwise."""
    g1, g2, g12 = float(p_f[14]), float(p_f[15]), float(p_f[16])
    if i == 0 and j == 0:
        return (2.0 * g1, -2.0 * g12, -2.0 * g12, 2.0 * g2)[2 * a + b]
    if i == 1 and j == 1:
        return (2.0 * g1, 2.0 * g12, 2.0 * g12, 2.0 * g2)[2 * a + b]
    return 0.0


def _d2(D2, i, j, a):
    """Reference idx_d2(i, j, a): xx for (0,0), yy for (1,1), mixed otherwise."""
    if i == j:
        return D2[a, i]
    return D2[a, 2]


def vacuum_energy(p_f):
    """Constant subtracted per site (kernels.py:168)."""
    a1, a2 = float(p_f[8]), float(p_f[9])
    b1, b2, b3, b4 = (float(p_f[k]) for k in (10, 11, 12, 13))
    u1, u2, th = float(p_f[17]), float(p_f[18]), float(p_f[19])
    return (a1 * u1 * u1 + a2 * u2 * u2 + b1 * u1 * u1 * u1 * u1 + b2 * u2 * u2 * u2 * u2
            + b3 * u1 * u1 * u2 * u2 + 2.0 * b4 * u1 * u1 * u2 * u2 * math.cos(2.0 * th))


def energy_density(F, D1, p_i, p_f):
    """kernels.py:96-170."""
    kappa, Q = float(p_f[6]), float(p_f[7])
    a1, a2 = float(p_f[8]), float(p_f[9])
    b1, b2, b3, b4 = (float(p_f[k]) for k in (10, 11, 12, 13))
    gv = float(p_f[4])
    A = (F[0], F[1])
    p1r, p1i, p2r, p2i = F[2], F[3], F[4], F[5]
    s1 = p1r * p1r + p1i * p1i
    s2 = p2r * p2r + p2i * p2i
    F_mag = 0.5 * kappa * kappa * (D1[1, 0] - D1[0, 1]) ** 2
    F_pot = a1 * s1 + a2 * s2
    F_pot = F_pot + b1 * s1 * s1 + b2 * s2 * s2 + b3 * s1 * s2
    F_pot = F_pot + 2.0 * b4 * (p1r * p1r * p2r * p2r + p1i * p1i * p2i * p2i - p1i * p1i * p2r * p2r
                                - p1r * p1r * p2i * p2i + 4.0 * p1r * p1i * p2r * p2i)
    F_kin = 0.0 * p1r
    for i in range(2):
        Ai = A[i]
        for j in range(2):
            Aj = A[j]
            g00, g01, g10, g11 = (gamma_entry(i, j, a, b, p_f) for a, b in ((0, 0), (0, 1), (1, 0), (1, 1)))
            d1r_i, d1r_j = D1[2, i], D1[2, j]
            d1i_i, d1i_j = D1[3, i], D1[3, j]
            d2r_i, d2r_j = D1[4, i], D1[4, j]
            d2i_i, d2i_j = D1[5, i], D1[5, j]
            F_kin = F_kin + 0.5 * g00 * (d1r_i * d1r_j + d1i_i * d1i_j + Q * Ai * (p1r * d1i_j - p1i * d1r_j)
                                         + Q * Aj * (p1r * d1i_i - p1i * d1r_i) + Q * Q * Ai * Aj * s1)
            F_kin = F_kin + 0.5 * g01 * (d1r_i * d2r_j + d1i_i * d2i_j + Q * Ai * (p1r * d2i_j - p1i * d2r_j)
                                         + Q * Aj * (p2r * d1i_i - p2i * d1r_i)
                                         + Q * Q * Ai * Aj * (p1r * p2r + p1i * p2i))
            F_kin = F_kin + 0.5 * g10 * (d2r_i * d1r_j + d2i_i * d1i_j + Q * Ai * (p2r * d1i_j - p2i * d1r_j)
                                         + Q * Aj * (p1r * d2i_i - p1i * d2r_i)
                                         + Q * Q * Ai * Aj * (p2r * p1r + p2i * p1i))
            F_kin = F_kin + 0.5 * g11 * (d2r_i * d2r_j + d2i_i * d2i_j + Q * Ai * (p2r * d2i_j - p2i * d2r_j)
                                         + Q * Aj * (p2r * d2i_i - p2i * d2r_i) + Q * Q * Ai * Aj * s2)
    e = F_kin + F_mag + F_pot
    e = e - vacuum_energy(p_f)
    return e * gv


def variation(F, D1, D2, p_i, p_f):
    """kernels.py:555-658; no mask, no projection."""
    kappa, Q = float(p_f[6]), float(p_f[7])
    a1, a2 = float(p_f[8]), float(p_f[9])
    b1, b2, b3, b4 = (float(p_f[k]) for k in (10, 11, 12, 13))
    A = (F[0], F[1])
    p1r, p1i, p2r, p2i = F[2], F[3], F[4], F[5]
    EG = [np.zeros_like(p1r) for _ in range(6)]
    for j in range(2):
        for k in range(2):
            Ak = A[k]
            g00, g01, g10, g11 = (gamma_entry(k, j, a, b, p_f) for a, b in ((0, 0), (0, 1), (1, 0), (1, 1)))
            val = kappa * kappa * (_d2(D2, j, k, k) - _d2(D2, k, k, j))
            val = val - Q * g00 * (p1i * D1[2, k] - p1r * D1[3, k] - Q * Ak * (p1r * p1r + p1i * p1i))
            val = val - Q * g01 * (p2i * D1[2, k] - p2r * D1[3, k] - Q * Ak * (p1r * p2r + p1i * p2i))
            val = val - Q * g10 * (p1i * D1[4, k] - p1r * D1[5, k] - Q * Ak * (p2r * p1r + p2i * p1i))
            val = val - Q * g11 * (p2i * D1[4, k] - p2r * D1[5, k] - Q * Ak * (p2r * p2r + p2i * p2i))
            EG[j] = EG[j] + val
    s1 = p1r * p1r + p1i * p1i
    s2 = p2r * p2r + p2i * p2i
    d2 = p2r * p2r - p2i * p2i
    d1 = p1r * p1r - p1i * p1i
    EG[2] = EG[2] + 2.0 * (2.0 * a1 * p1r + 4.0 * b1 * p1r * s1 + 2.0 * b3 * p1r * s2
                           + 2.0 * b4 * (2.0 * p1r * d2 + 4.0 * p1i * p2r * p2i))
    EG[3] = EG[3] + 2.0 * (2.0 * a1 * p1i + 4.0 * b1 * p1i * s1 + 2.0 * b3 * p1i * s2
                           + 2.0 * b4 * (-2.0 * p1i * d2 + 4.0 * p1r * p2r * p2i))
    EG[4] = EG[4] + 2.0 * (2.0 * a2 * p2r + 4.0 * b2 * p2r * s2 + 2.0 * b3 * p2r * s1
                           + 2.0 * b4 * (2.0 * p2r * d1 + 4.0 * p1r * p1i * p2i))
    EG[5] = EG[5] + 2.0 * (2.0 * a2 * p2i + 4.0 * b2 * p2i * s2 + 2.0 * b3 * p2i * s1
                           + 2.0 * b4 * (-2.0 * p2i * d1 + 4.0 * p1r * p1i * p2r))
    for i in range(2):
        Ai = A[i]
        for j in range(2):
            Aj = A[j]
            g00, g01, g10, g11 = (gamma_entry(i, j, a, b, p_f) for a, b in ((0, 0), (0, 1), (1, 0), (1, 1)))
            dA = D1[j, i]   # reference d1fd1x[idx_d1(i, j)] = d_i A_j
            t2 = _d2(D2, i, j, 2) - Q * Q * Ai * Aj * p1r - Q * (p1i * dA + Ai * D1[3, j] + Aj * D1[3, i])
            t4 = _d2(D2, i, j, 4) - Q * Q * Ai * Aj * p2r - Q * (p2i * dA + Ai * D1[5, j] + Aj * D1[5, i])
            t3 = _d2(D2, i, j, 3) - Q * Q * Ai * Aj * p1i + Q * (p1r * dA + Ai * D1[2, j] + Aj * D1[2, i])
            t5 = _d2(D2, i, j, 5) - Q * Q * Ai * Aj * p2i + Q * (p2r * dA + Ai * D1[4, j] + Aj * D1[4, i])
            EG[2] = EG[2] - g00 * t2
            EG[2] = EG[2] - g01 * t4
            EG[3] = EG[3] - g00 * t3
            EG[3] = EG[3] - g01 * t5
            EG[4] = EG[4] - g10 * t2
            EG[4] = EG[4] - g11 * t4
            EG[5] = EG[5] - g10 * t3
            EG[5] = EG[5] - g11 * t5
    return np.stack(EG)


def vortex_density(F, D1, p_i, p_f):
    """``(d_x A2 - d_y A1) * gv / (2 pi Q)`` (kernels.py:207-240)."""
    gv, q = float(p_f[4]), float(p_f[7])
    return (D1[1, 0] - D1[0, 1]) * (gv / (2.0 * math.pi * q))


CHARGES = {"vortex_number": vortex_density}
