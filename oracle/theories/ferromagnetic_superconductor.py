"""Ferromagnetic superconductor (oracle, FP64).

TEST INFRASTRUCTURE ONLY.  Restates soliton_solver/theories/
ferromagnetic_superconductor/kernels.py.  Fields: m1, m2, m3, A1, A2, A3,
psi1, psi2.  Packed parameters (ferromagnetic_superconductor/params.py:266-312):
p_f[6] q, [7] ha, [8] hb, [9] eta1, [10] eta2, [11] u1, [12] vortex_number,
[13] ainf, [14] alpha, [15] beta, [16] gamma, [17] M0, [18] skN, [19] rotation.
The magnetisation is constrained to |m| = M0 (kernels.py:46-120).
"""
from __future__ import annotations

import math

import numpy as np

NAME = "Ferromagnetic superconductor"
HAS_CONSTRAINT = True


def _pf(p_f):
    return {k: float(p_f[i]) for k, i in (("q", 6), ("ha", 7), ("hb", 8), ("eta1", 9), ("eta2", 10),
                                           ("alpha", 14), ("beta", 15), ("gamma", 16), ("M0", 17))}


def energy_density(F, D1, p_i, p_f):
    """kernels.py:124-181."""
    c = _pf(p_f)
    q, ha, hb, e1, e2, al, be, ga = c["q"], c["ha"], c["hb"], c["eta1"], c["eta2"], c["alpha"], c["beta"], c["gamma"]
    gv = float(p_f[4])
    m1, m2, m3, A1, A2, A3, p1, p2 = (F[k] for k in range(8))
    dm = [(D1[a, 0], D1[a, 1]) for a in range(3)]
    dA1dx, dA1dy = D1[3, 0], D1[3, 1]
    dA2dx, dA2dy = D1[4, 0], D1[4, 1]
    dA3dx, dA3dy = D1[5, 0], D1[5, 1]
    dp1dx, dp1dy = D1[6, 0], D1[6, 1]
    dp2dx, dp2dy = D1[7, 0], D1[7, 1]
    psq = p1 * p1 + p2 * p2
    msq = m1 * m1 + m2 * m2 + m3 * m3
    Asq = A1 * A1 + A2 * A2 + A3 * A3
    gradm = (dm[0][0] * dm[0][0] + dm[1][0] * dm[1][0] + dm[2][0] * dm[2][0]
             + dm[0][1] * dm[0][1] + dm[1][1] * dm[1][1] + dm[2][1] * dm[2][1])
    e = 0.0 * F[0]
    e = e + ha / 2.0 * psq
    e = e + hb / 4.0 * psq * psq
    e = e + 0.5 * dp1dx * dp1dx + 0.5 * dp1dy * dp1dy + 0.5 * dp2dx * dp2dx + 0.5 * dp2dy * dp2dy
    e = e + 0.5 * q * q * Asq * psq
    e = e + q * A1 * (p1 * dp2dx - p2 * dp1dx) + q * A2 * (p1 * dp2dy - p2 * dp1dy)
    e = e + 0.5 * dA2dx * dA2dx + 0.5 * dA1dy * dA1dy - dA2dx * dA1dy + 0.5 * dA3dx * dA3dx + 0.5 * dA3dy * dA3dy
    e = e + al / 2.0 * msq + be / 4.0 * msq * msq
    e = e + ga * ga / 2.0 * gradm
    e = e + m2 * dA3dx - m1 * dA3dy - m3 * dA2dx + m3 * dA1dy
    e = e + e1 * msq * psq
    e = e + e2 * gradm * psq
    denom = (16.0 * e1 * e1 - 4.0 * hb * be)
    e = e - ((al * al * hb + ha * ha * be - 4.0 * ha * al * e1) / denom)
    return e * gv


def variation(F, D1, D2, p_i, p_f):
    """kernels.py:533-619, then the projection against m/M0 (:616)."""
    c = _pf(p_f)
    q, ha, hb, e1, e2, al, be, ga, M0 = (c[k] for k in ("q", "ha", "hb", "eta1", "eta2", "alpha", "beta", "gamma", "M0"))
    m1, m2, m3, A1, A2, A3, p1, p2 = (F[k] for k in range(8))
    dm1dx, dm1dy = D1[0, 0], D1[0, 1]
    dm2dx, dm2dy = D1[1, 0], D1[1, 1]
    dm3dx, dm3dy = D1[2, 0], D1[2, 1]
    dA1dx, dA1dy = D1[3, 0], D1[3, 1]
    dA2dx, dA2dy = D1[4, 0], D1[4, 1]
    dA3dx, dA3dy = D1[5, 0], D1[5, 1]
    dp1dx, dp1dy = D1[6, 0], D1[6, 1]
    dp2dx, dp2dy = D1[7, 0], D1[7, 1]
    psq = p1 * p1 + p2 * p2
    msq = m1 * m1 + m2 * m2 + m3 * m3
    dmsq = dm1dx * dm1dx + dm2dx * dm2dx + dm3dx * dm3dx + dm1dy * dm1dy + dm2dy * dm2dy + dm3dy * dm3dy
    Asq = A1 * A1 + A2 * A2 + A3 * A3
    lap = [D2[a, 0] + D2[a, 1] for a in range(8)]
    g0 = al * m1 + be * m1 * msq - ga * ga * D2[0, 0] - ga * ga * D2[0, 1] - dA3dy + 2.0 * e1 * psq * m1 - 2.0 * e2 * lap[0] * psq
    g1 = al * m2 + be * m2 * msq - ga * ga * D2[1, 0] - ga * ga * D2[1, 1] + dA3dx + 2.0 * e1 * psq * m2 - 2.0 * e2 * lap[1] * psq
    g2 = (al * m3 + be * m3 * msq - ga * ga * D2[2, 0] - ga * ga * D2[2, 1] - dA2dx + dA1dy + 2.0 * e1 * psq * m3
          - 2.0 * e2 * lap[2] * psq)
    g3 = q * q * A1 * psq + q * (p1 * dp2dx - p2 * dp1dx) + D2[4, 2] - D2[3, 1] - dm3dy
    g4 = q * q * A2 * psq + q * (p1 * dp2dy - p2 * dp1dy) + D2[3, 2] - D2[4, 0] + dm3dx
    g5 = q * q * A3 * psq - D2[5, 0] - D2[5, 1] - dm2dx + dm1dy
    g6 = (ha * p1 + hb * p1 * psq - D2[6, 0] - D2[6, 1] + q * q * Asq * p1 + 2.0 * q * A1 * dp2dx + 2.0 * q * A2 * dp2dy
          + q * p2 * (dA1dx + dA2dy) + 2.0 * e1 * msq * p1 + 2.0 * e2 * p1 * p1 * dmsq)
    g7 = (ha * p2 + hb * p2 * psq - D2[7, 0] - D2[7, 1] + q * q * Asq * p2 - 2.0 * q * A1 * dp1dx - 2.0 * q * A2 * dp1dy
          - q * p1 * (dA1dx + dA2dy) + 2.0 * e1 * msq * p2 + 2.0 * e2 * p2 * p2 * dmsq)
    EG = np.stack([g0, g1, g2, g3, g4, g5, g6, g7])
    project(EG, F, p_i, p_f)
    return EG


def normalize(F, p_i, p_f):
    """|m| = M0 (kernels.py:46-83): m *= M0/|m|, sites with |m| = 0 unchanged."""
    M0 = float(p_f[17])
    s = np.sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2])
    nz = s != 0.0
    for a in range(3):
        F[a] = np.where(nz, F[a] * (M0 / np.where(nz, s, 1.0)), F[a])


def project(G, F, p_i, p_f):
    """G -= (G.m/M0) m/M0 (kernels.py:85-120)."""
    M0 = float(p_f[17])
    lm = 0.0 * F[0]
    for a in range(3):
        lm = lm + G[a] * F[a] / M0
    for a in range(3):
        G[a] = G[a] - lm * F[a] / M0


def skyrmion_density(F, D1, p_i, p_f):
    """kernels.py:218-253, normalised by M0^3."""
    gv, M0 = float(p_f[4]), float(p_f[17])
    m1, m2, m3 = F[0], F[1], F[2]
    cx0 = D1[1, 0] * D1[2, 1] - D1[2, 0] * D1[1, 1]
    cx1 = D1[2, 0] * D1[0, 1] - D1[0, 0] * D1[2, 1]
    cx2 = D1[0, 0] * D1[1, 1] - D1[1, 0] * D1[0, 1]
    return (m1 * cx0 + m2 * cx1 + m3 * cx2) * (gv / (4.0 * math.pi * M0 * M0 * M0))


def vortex_density(F, D1, p_i, p_f, which=2):
    """kernels.py:290-329: which 0 d_y A3, 1 -d_x A3, else d_x A2 - d_y A1; over 2 pi q."""
    gv, q = float(p_f[4]), float(p_f[6])
    if which == 0:
        b = D1[5, 1]
    elif which == 1:
        b = -D1[5, 0]
    else:
        b = D1[4, 0] - D1[3, 1]
    return b * (gv / (2.0 * math.pi * q))


CHARGES = {"skyrmion_number": skyrmion_density, "vortex_number": vortex_density}
