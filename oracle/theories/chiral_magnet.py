"""Chiral magnet: energy density, variation, charges (oracle, FP64).

TEST INFRASTRUCTURE ONLY.  Restates soliton_solver/theories/chiral_magnet/
kernels.py of the reference.  Packed parameters (chiral_magnet/params.py:333-378):
p_i[10] nmag, [11] dresselhaus, [12] rashba, [13] heusler, [14] hybrid,
[15] demag; p_f[6] coup_K, [7] coup_h, [8] coup_mu.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import core

NAME = "Chiral magnet"
HAS_CONSTRAINT = True


def levi(i, j, k):
    """Levi-Civita symbol as the reference's inner ``levi`` (kernels.py:162-169)."""
    if (i, j, k) in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        return 1.0
    if (i, j, k) in ((0, 2, 1), (2, 1, 0), (1, 0, 2)):
        return -1.0
    return 0.0


def dmi_matrix(p_i):
    """(D00, D01, D10, D11) by the first set flag (kernels.py:171-183)."""
    D00 = D01 = D10 = D11 = 0.0
    if int(p_i[11]):
        D00, D11 = -1.0, -1.0
    elif int(p_i[12]):
        D01, D10 = -1.0, 1.0
    elif int(p_i[13]):
        D00, D11 = 1.0, -1.0
    elif int(p_i[14]):
        s = 0.7071067811865476
        D00, D01, D10, D11 = -s, -s, s, -s
    return D00, D01, D10, D11


def energy_density(F, D1, p_i, p_f):
    """Per-site energy times gv (kernels.py:124-216)."""
    gv = float(p_f[4])
    coup_K, coup_h, coup_mu = float(p_f[6]), float(p_f[7]), float(p_f[8])
    nc, nmag, demag = int(p_i[3]), int(p_i[10]), int(p_i[15])
    D00, D01, D10, D11 = dmi_matrix(p_i)
    e = 0.0 * F[0]
    for a in range(nmag):
        for i in range(nc):
            d = D1[a, i]
            e = e + 0.5 * d * d
    for i in range(nc):
        for j in range(nc):
            if j == 0 and i == 0:
                Dij = D00
            elif j == 0 and i == 1:
                Dij = D01
            elif j == 1 and i == 0:
                Dij = D10
            else:
                Dij = D11
            if Dij == 0.0:
                continue
            for k in range(nmag):
                mk = F[k]
                for l in range(nmag):
                    eps = levi(j, k, l)
                    if eps != 0.0:
                        e = e + (Dij * eps * mk * D1[l, i])
        if demag:
            e = e - (F[3] * D1[i, i] + D1[3, i] * D1[3, i] / (2.0 * coup_mu))
    mz = F[nmag - 1]
    e = e + coup_h * (1.0 - mz)
    e = e + coup_K * (1.0 - mz * mz)
    return e * gv


def variation(F, D1, D2, p_i, p_f):
    """Reference energy gradient EG (kernels.py:397-511), before ``l = -dt EG``.

    Includes the boundary-band mask (:505-507) and the tangent projection
    (:509) with the stage field as given (no renormalisation).
    """
    coup_K, coup_h, coup_mu = float(p_f[6]), float(p_f[7]), float(p_f[8])
    nc, nf, nmag, demag = int(p_i[3]), int(p_i[4]), int(p_i[10]), int(p_i[15])
    D00, D01, D10, D11 = dmi_matrix(p_i)
    EG = np.zeros_like(F)
    for i in range(nmag):
        lap = 0.0
        for j in range(nc):
            lap = lap + D2[i, j]
        EG[i] = EG[i] - lap
    if D00 != 0.0 or D01 != 0.0 or D10 != 0.0 or D11 != 0.0:
        for i in range(nmag):
            for j in range(nmag):
                eps = levi(0, i, j)
                if eps != 0.0:
                    EG[i] = EG[i] + 2.0 * D00 * eps * D1[j, 0]
                eps = levi(0, i, j)
                if eps != 0.0 and D01 != 0.0:
                    EG[i] = EG[i] + 2.0 * D01 * eps * D1[j, 1]
                eps = levi(1, i, j)
                if eps != 0.0 and D10 != 0.0:
                    EG[i] = EG[i] + 2.0 * D10 * eps * D1[j, 0]
                eps = levi(1, i, j)
                if eps != 0.0 and D11 != 0.0:
                    EG[i] = EG[i] + 2.0 * D11 * eps * D1[j, 1]
    last = nmag - 1
    mz = F[last]
    EG[last] = EG[last] - coup_h
    EG[last] = EG[last] - 2.0 * coup_K * mz
    if demag:
        for i in range(nc):
            EG[i] = EG[i] + D1[3, i]
            EG[3] = EG[3] + (coup_mu * D1[i, i] - D2[3, i])
    EG[:, core.band_mask(p_i)] = 0.0
    core.tangent_project(EG, F, nmag)
    return EG


def normalize(F, p_i, p_f):
    core.unit_norm(F, int(p_i[10]))


def project(V, F, p_i, p_f):
    core.tangent_project(V, F, int(p_i[10]))


def skyrmion_density(F, D1, p_i, p_f):
    """``m . (d_x m x d_y m) * gv / (4 pi)`` (kernels.py:253-291)."""
    gv = float(p_f[4])
    m1, m2, m3 = F[0], F[1], F[2]
    dm1dx, dm2dx, dm3dx = D1[0, 0], D1[1, 0], D1[2, 0]
    dm1dy, dm2dy, dm3dy = D1[0, 1], D1[1, 1], D1[2, 1]
    cx0 = dm2dx * dm3dy - dm3dx * dm2dy
    cx1 = dm3dx * dm1dy - dm1dx * dm3dy
    cx2 = dm1dx * dm2dy - dm2dx * dm1dy
    charge = m1 * cx0 + m2 * cx1 + m3 * cx2
    return charge * (gv / (4.0 * math.pi))


def magnetic_charge_density(F, D1, p_i, p_f):
    """``-(d_x m1 + d_y m2) * gv`` (kernels.py:328-360)."""
    gv = float(p_f[4])
    charge = np.zeros(F.shape[1:])
    for i in range(int(p_i[3])):
        charge = charge - D1[i, i]
    return charge * gv


CHARGES = {
    "skyrmion_number": skyrmion_density,
    "magnetic_charge": magnetic_charge_density,
}
