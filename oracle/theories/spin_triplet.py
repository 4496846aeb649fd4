"""Spin-triplet superconducting ferromagnet (oracle, FP64).

TEST INFRASTRUCTURE ONLY.  Restates soliton_solver/theories/
spin_triplet_superconducting_magnet/kernels.py.  Fields: m1, m2, m3, A1, A2, A3,
psi1 (re, im), psi2 (re, im).  Packed parameters
(spin_triplet_superconducting_magnet/params.py, pack_device_params):
p_f[6] q, [7] ha, [8] hb1, [9] hb2, [10] hc, [11] u1, [12] u2, [13] vortex1_number,
[14] vortex2_number, [15] vortex_number, [16] ainf, [17] alpha, [18] beta,
[19] gamma, [20] M0, [21] skyrmion_number, [22] skyrmion_rotation, [23..26] ansatz
flags.  The magnetisation is constrained to |m| = M0 (kernels.py:47-120).
"""
from __future__ import annotations

import math

import numpy as np

NAME = "Spin triplet superconducting ferromagnet"
HAS_CONSTRAINT = True


def _c(p_f):
    return [float(p_f[i]) for i in (6, 7, 8, 9, 10, 11, 12, 17, 18, 19, 20)]


def energy_density(F, D1, p_i, p_f):
    """kernels.py:125-191."""
    q, ha, hb1, hb2, hc, u1, u2, al, be, ga, M0 = _c(p_f)
    gv = float(p_f[4])
    m1, m2, m3, A1, A2, A3, p1, p2, p3, p4 = (F[k] for k in range(10))
    d = [(D1[a, 0], D1[a, 1]) for a in range(10)]
    p12 = p1 * p1 + p2 * p2
    p34 = p3 * p3 + p4 * p4
    Asq = A1 * A1 + A2 * A2 + A3 * A3
    msq = m1 * m1 + m2 * m2 + m3 * m3
    psq = p12 + p34
    e = 0.0 * F[0]
    e = e + ha / 2.0 * (p1 * p1 + p2 * p2 + p3 * p3 + p4 * p4)
    e = e + hb1 / 4.0 * p12 * p12 + hb1 / 4.0 * p34 * p34
    e = e + hb2 * p12 * p34
    e = e + 2.0 * hc * (p1 * p3 + p2 * p4)
    for a in (6, 7, 8, 9):
        e = e + 0.5 * d[a][0] * d[a][0] + 0.5 * d[a][1] * d[a][1]
    e = e + 0.5 * q * q * Asq * psq
    e = e + q * A1 * (p1 * d[7][0] - p2 * d[6][0] + p3 * d[9][0] - p4 * d[8][0])
    e = e + q * A2 * (p1 * d[7][1] - p2 * d[6][1] + p3 * d[9][1] - p4 * d[8][1])
    e = e + 0.5 * d[4][0] * d[4][0] + 0.5 * d[3][1] * d[3][1] - d[4][0] * d[3][1]
    e = e + 0.5 * d[5][0] * d[5][0] + 0.5 * d[5][1] * d[5][1]
    e = e + al / 2.0 * msq + be / 4.0 * msq * msq
    e = e + 0.5 * ga * ga * sum(d[a][0] * d[a][0] + d[a][1] * d[a][1] for a in range(3))
    e = e + m2 * d[5][0] - m1 * d[5][1] - m3 * d[4][0] + m3 * d[3][1]
    e = e - (ha / 2.0 * (u1 * u1 + u2 * u2) + hb1 / 4.0 * (u1 ** 4 + u2 ** 4) + hb2 * u1 * u1 * u2 * u2
             + 2.0 * hc * u1 * u2 + al / 2.0 * M0 * M0 + be / 4.0 * M0 ** 4)
    return e * gv


def variation(F, D1, D2, p_i, p_f):
    """kernels.py:601-690 (g0..g9), then the projection against m/M0 (:687)."""
    q, ha, hb1, hb2, hc, u1, u2, al, be, ga, M0 = _c(p_f)
    m1, m2, m3, A1, A2, A3, p1, p2, p3, p4 = (F[k] for k in range(10))
    d = [(D1[a, 0], D1[a, 1]) for a in range(10)]
    msq = m1 * m1 + m2 * m2 + m3 * m3
    Asq = A1 * A1 + A2 * A2 + A3 * A3
    p12 = p1 * p1 + p2 * p2
    p34 = p3 * p3 + p4 * p4
    psq = p12 + p34
    g = [None] * 10
    g[0] = al * m1 + be * m1 * msq - ga * ga * D2[0, 0] - ga * ga * D2[0, 1] - d[5][1]
    g[1] = al * m2 + be * m2 * msq - ga * ga * D2[1, 0] - ga * ga * D2[1, 1] + d[5][0]
    g[2] = al * m3 + be * m3 * msq - ga * ga * D2[2, 0] - ga * ga * D2[2, 1] - d[4][0] + d[3][1]
    g[3] = (q * q * A1 * psq + q * (p1 * d[7][0] - p2 * d[6][0] + p3 * d[9][0] - p4 * d[8][0]) + D2[4, 2] - D2[3, 1]
            - d[2][1])
    g[4] = (q * q * A2 * psq + q * (p1 * d[7][1] - p2 * d[6][1] + p3 * d[9][1] - p4 * d[8][1]) + D2[3, 2] - D2[4, 0]
            + d[2][0])
    g[5] = q * q * A3 * psq - D2[5, 0] - D2[5, 1] - d[1][0] + d[0][1]
    divA = d[3][0] + d[4][1]
    g[6] = (ha * p1 + hb1 * p1 * p12 - D2[6, 0] - D2[6, 1] + q * q * Asq * p1 + 2.0 * q * A1 * d[7][0]
            + 2.0 * q * A2 * d[7][1] + q * p2 * divA + 2.0 * hb2 * p1 * p34 + 2.0 * hc * p3)
    g[7] = (ha * p2 + hb1 * p2 * p12 - D2[7, 0] - D2[7, 1] + q * q * Asq * p2 - 2.0 * q * A1 * d[6][0]
            - 2.0 * q * A2 * d[6][1] - q * p1 * divA + 2.0 * hb2 * p2 * p34 + 2.0 * hc * p4)
    g[8] = (ha * p3 + hb1 * p3 * p34 - D2[8, 0] - D2[8, 1] + q * q * Asq * p3 + 2.0 * q * A1 * d[9][0]
            + 2.0 * q * A2 * d[9][1] + q * p4 * divA + 2.0 * hb2 * p3 * p12 + 2.0 * hc * p1)
    g[9] = (ha * p4 + hb1 * p4 * p34 - D2[9, 0] - D2[9, 1] + q * q * Asq * p4 - 2.0 * q * A1 * d[8][0]
            - 2.0 * q * A2 * d[8][1] - q * p3 * divA + 2.0 * hb2 * p4 * p12 + 2.0 * hc * p2)
    EG = np.stack(g)
    project(EG, F, p_i, p_f)
    return EG


def normalize(F, p_i, p_f):
    """|m| = M0 (kernels.py:47-84)."""
    M0 = float(p_f[20])
    s = np.sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2])
    nz = s != 0.0
    for a in range(3):
        F[a] = np.where(nz, F[a] * (M0 / np.where(nz, s, 1.0)), F[a])


def project(G, F, p_i, p_f):
    """G -= (G.m/M0) m/M0 (kernels.py:86-120)."""
    M0 = float(p_f[20])
    lm = 0.0 * F[0]
    for a in range(3):
        lm = lm + G[a] * F[a] / M0
    for a in range(3):
        G[a] = G[a] - lm * F[a] / M0


def skyrmion_density(F, D1, p_i, p_f):
    """kernels.py:229-263, normalised by M0^3."""
    gv, M0 = float(p_f[4]), float(p_f[20])
    cx0 = D1[1, 0] * D1[2, 1] - D1[2, 0] * D1[1, 1]
    cx1 = D1[2, 0] * D1[0, 1] - D1[0, 0] * D1[2, 1]
    cx2 = D1[0, 0] * D1[1, 1] - D1[1, 0] * D1[0, 1]
    return (F[0] * cx0 + F[1] * cx1 + F[2] * cx2) * (gv / (4.0 * math.pi * M0 * M0 * M0))


def vortex_density(F, D1, p_i, p_f, which=2):
    """kernels.py:301-339: which 0 d_y A3, 1 -d_x A3, else d_x A2 - d_y A1; over 2 pi q."""
    gv, q = float(p_f[4]), float(p_f[6])
    if which == 0:
        b = D1[5, 1]
    elif which == 1:
        b = -D1[5, 0]
    else:
        b = D1[4, 0] - D1[3, 1]
    return b * (gv / (2.0 * math.pi * q))


CHARGES = {"skyrmion_number": skyrmion_density, "vortex_number": vortex_density}
