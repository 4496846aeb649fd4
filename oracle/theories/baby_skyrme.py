"""Baby Skyrme model (oracle, FP64).

TEST INFRASTRUCTURE ONLY.  Restates soliton_solver/theories/baby_skyrme/
kernels.py.  Packed parameters (baby_skyrme/params.py:331-378):
p_i[10] nmag, [11] standard, [12] holomorphic, [13] easy plane,
[14] dihedral2, [15] aloof, [16] dihedral3, [17] broken, [18] double vacua,
[19] N; p_f[6] mpi, [7] kappa, [8] skyrmion_number.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import core

NAME = "Baby Skyrme model"
HAS_CONSTRAINT = True


def energy_density(F, D1, p_i, p_f):
    """kernels.py:126-199."""
    gv = float(p_f[4])
    mpi2 = float(p_f[6]) * float(p_f[6])
    kappa2 = float(p_f[7]) * float(p_f[7])
    m1, m2, m3 = F[0], F[1], F[2]
    dm1dx, dm1dy = D1[0, 0], D1[0, 1]
    dm2dx, dm2dy = D1[1, 0], D1[1, 1]
    dm3dx, dm3dy = D1[2, 0], D1[2, 1]
    gx = dm1dx * dm1dx + dm2dx * dm2dx + dm3dx * dm3dx
    gy = dm1dy * dm1dy + dm2dy * dm2dy + dm3dy * dm3dy
    e = 0.5 * (gx + gy)
    cx = dm2dx * dm3dy - dm3dx * dm2dy
    cy = dm3dx * dm1dy - dm1dx * dm3dy
    cz = dm1dx * dm2dy - dm2dx * dm1dy
    e = e + 0.5 * kappa2 * (cx * cx + cy * cy + cz * cz)
    if int(p_i[11]):
        e = e + mpi2 * (1.0 - m3)
    if int(p_i[12]):
        t = 1.0 - m3
        e = e + mpi2 * t * t * t * t
    if int(p_i[13]):
        e = e + 0.5 * mpi2 * m1 * m1
    if int(p_i[14]):
        e = e + 0.5 * mpi2 * (1.0 - m1 * m1) * (1.0 - m3 * m3)
    if int(p_i[15]):
        t = 1.0 - m3
        e = e + 0.5 * mpi2 * t * (1.0 + t * t * t)
    if int(p_i[16]):
        e = e + (16.0 * mpi2 * (1.0 - m3) * (1.0 + 3.0 * m3 * m3 + 3.0 * m1 * m2 * m2 - m1 * m1 * m1))
    if int(p_i[17]):
        N = int(p_i[19])
        zr, zi = m1, m2
        for _ in range(N - 1):
            zr, zi = zr * m1 - zi * m2, zr * m2 + zi * m1
        wr, wi = 1.0 - zr, -zi
        e = e + mpi2 * (1.0 - m3) * (wr * wr + wi * wi)
    if int(p_i[18]):
        e = e + mpi2 * (1.0 - m3 * m3)
    return e * gv


def variation(F, D1, D2, p_i, p_f):
    """kernels.py:307-458: exchange + Skyrme-term Euler-Lagrange expression,
    potential derivatives, then the tangent projection (no band mask)."""
    mpi2 = float(p_f[6]) * float(p_f[6])
    kappa2 = float(p_f[7]) * float(p_f[7])
    m1, m2, m3 = F[0], F[1], F[2]
    F0 = [D1[0, 0], D1[1, 0], D1[2, 0]]   # d/dx of m_a
    F1 = [D1[0, 1], D1[1, 1], D1[2, 1]]   # d/dy of m_a
    Sxx = [D2[0, 0], D2[1, 0], D2[2, 0]]
    Syy = [D2[0, 1], D2[1, 1], D2[2, 1]]
    Sxy = [D2[0, 2], D2[1, 2], D2[2, 2]]  # both off-diagonal slots hold the mixed entry
    s0 = F0[0] * F0[0] + F0[1] * F0[1] + F0[2] * F0[2]
    s1 = F1[0] * F1[0] + F1[1] * F1[1] + F1[2] * F1[2]
    sall = s0 + s1
    g = [np.zeros_like(m1) for _ in range(3)]
    for a in range(3):
        g[a] = g[a] - (Sxx[a] + Syy[a])
    for a in range(3):
        g[a] = g[a] - kappa2 * Sxx[a] * sall
        g[a] = g[a] + kappa2 * Sxx[a] * s0
    A = [Sxy[a] * F1[a] - Syy[a] * F0[a] for a in range(3)]
    sumA = A[0] + A[1] + A[2]
    dot01 = F0[0] * F1[0] + F0[1] * F1[1] + F0[2] * F1[2]
    for a in range(3):
        g[a] = g[a] - kappa2 * (F0[a] * sumA - Sxy[a] * dot01)
    for a in range(3):
        g[a] = g[a] - kappa2 * Syy[a] * sall
        g[a] = g[a] + kappa2 * Syy[a] * s1
    B = [Sxy[a] * F0[a] - Sxx[a] * F1[a] for a in range(3)]
    sumB = B[0] + B[1] + B[2]
    for a in range(3):
        g[a] = g[a] - kappa2 * (F1[a] * sumB - Sxy[a] * dot01)
    if int(p_i[11]):
        g[2] = g[2] - mpi2
    if int(p_i[12]):
        t = 1.0 - m3
        g[2] = g[2] - 4.0 * mpi2 * t * t * t
    if int(p_i[13]):
        g[0] = g[0] + mpi2 * m1
    if int(p_i[14]):
        g[0] = g[0] - mpi2 * (1.0 - m3 * m3) * m1
        g[2] = g[2] - mpi2 * (1.0 - m1 * m1) * m3
    if int(p_i[15]):
        t = 1.0 - m3
        t3 = t * t * t
        g[2] = g[2] - mpi2 * (0.5 * (1.0 + t3) + 1.5 * t3)
    if int(p_i[16]):
        g[0] = g[0] + 48.0 * mpi2 * ((m2 * m2) - (m1 * m1)) * (1.0 - m3)
        g[1] = g[1] + 96.0 * mpi2 * m1 * m2 * (1.0 - m3)
        g[2] = g[2] - 16.0 * mpi2 * (-(m1 * m1 * m1) + 3.0 * m1 * (m2 * m2) + (1.0 - 3.0 * m3) * (1.0 - 3.0 * m3))
    if int(p_i[17]):
        N = int(p_i[19])
        r2 = m1 * m1 + m2 * m2
        theta = np.arctan2(m2, m1)
        rN2 = r2 ** (0.5 * N)
        rN = r2 ** (1.0 * N)
        c = np.cos(N * theta)
        s = np.sin(N * theta)
        norm1mzN = 1.0 - 2.0 * c * rN2 + rN
        denom = np.sqrt(norm1mzN)
        ok = (denom != 0.0) & (r2 != 0.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            pref = 2.0 * mpi2 * N * np.sqrt(norm1mzN) * (r2 ** (-1.0 + 0.5 * N)) * (-1.0 + m3) / denom
        g[0] = g[0] + np.where(ok, pref * (s * m2 + m1 * (c - rN2)), 0.0)
        g[1] = g[1] + np.where(ok, (-pref) * (s * m1 + m2 * (-c + rN2)), 0.0)
        g[2] = g[2] - mpi2 * norm1mzN
    if int(p_i[18]):
        g[2] = g[2] - 2.0 * mpi2 * m3
    EG = np.stack(g)
    core.tangent_project(EG, F, int(p_i[10]))
    return EG


def normalize(F, p_i, p_f):
    core.unit_norm(F, int(p_i[10]))


def project(V, F, p_i, p_f):
    core.tangent_project(V, F, int(p_i[10]))


def skyrmion_density(F, D1, p_i, p_f):
    """kernels.py:236-270."""
    gv = float(p_f[4])
    m1, m2, m3 = F[0], F[1], F[2]
    dm1dx, dm2dx, dm3dx = D1[0, 0], D1[1, 0], D1[2, 0]
    dm1dy, dm2dy, dm3dy = D1[0, 1], D1[1, 1], D1[2, 1]
    cx0 = dm2dx * dm3dy - dm3dx * dm2dy
    cx1 = dm3dx * dm1dy - dm1dx * dm3dy
    cx2 = dm1dx * dm2dy - dm2dx * dm1dy
    return (m1 * cx0 + m2 * cx1 + m3 * cx2) * (gv / (4.0 * math.pi))


CHARGES = {"skyrmion_number": skyrmion_density}
