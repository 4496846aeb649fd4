"""Chiral liquid crystal (oracle, FP64).

TEST INFRASTRUCTURE ONLY.  Restates soliton_solver/theories/liquid_crystal/
kernels.py.  Fields: n1, n2, n3 (director, |n| = 1) and Phi (electric
potential of the flexoelectric depolarisation).  Packed parameters
(liquid_crystal/params.py, pack_device_params): p_i[10] nmag, [11] dresselhaus
(twist), [12] rashba (splay-bend), [13] depol; p_f[6] coup_PotE, [7] coup_Potw0,
[8] coup_eps, [9] e1, [10] e3, [11] skyrmion_number, [12] skyrmion_rotation,
[13..16] ansatz flags.
"""
from __future__ import annotations

import math

import numpy as np

from oracle import core
from oracle.theories.chiral_magnet import levi

NAME = "Liquid crystal"
HAS_CONSTRAINT = True


def dmi_matrix(p_i):
    """(D00, D01, D10, D11): Dresselhaus (-1, 0, 0, -1) or else Rashba (0, -1, 1, 0) (kernels.py:171-177)."""
    if int(p_i[11]):
        return -1.0, 0.0, 0.0, -1.0
    if int(p_i[12]):
        return 0.0, -1.0, 1.0, 0.0
    return 0.0, 0.0, 0.0, 0.0


def _d2(D2, a, i, j):
    """d_i d_j of field a (the mixed entry is shared by (0, 1) and (1, 0))."""
    return D2[a, i] if i == j else D2[a, 2]


def energy_density(F, D1, p_i, p_f):
    """kernels.py:124-216."""
    gv = float(p_f[4])
    PotE, Potw0, e1, e3 = float(p_f[6]), float(p_f[7]), float(p_f[9]), float(p_f[10])
    nc, nmag, depol = int(p_i[3]), int(p_i[10]), int(p_i[13])
    D = dmi_matrix(p_i)
    e = 0.0 * F[0]
    for a in range(nmag):
        for i in range(nc):
            e = e + 0.5 * D1[a, i] * D1[a, i]
    for i in range(nc):
        for j in range(nc):
            Dij = D[0] if (i == 0 and j == 0) else D[1] if (i == 1 and j == 0) else D[2] if (i == 0 and j == 1) else D[3]
            if Dij == 0.0:
                continue
            for k in range(nmag):
                for l in range(nmag):
                    eps = levi(j, k, l)
                    if eps != 0.0:
                        e = e + Dij * eps * F[k] * D1[l, i]
    if depol:
        r = e3 / e1
        for i in range(nc):
            for j in range(nmag):
                e = e + r * F[j] * D1[j, i] * D1[3, i] / 2.0
            for j in range(nc):
                e = e + D1[i, i] * F[j] * D1[3, j] / 2.0
                e = e - r * F[j] * D1[i, j] * D1[3, i] / 2.0
    mz = F[nmag - 1]
    e = e + 0.5 * PotE * (1.0 - mz * mz)
    e = e + 0.5 * Potw0 * (1.0 - mz * mz)
    return e * gv


def variation(F, D1, D2, p_i, p_f):
    """kernels.py:403-509: exchange, DMI, potential, depolarisation (:492-500),
    boundary-band mask (:502-504), tangent projection (:506)."""
    PotE, Potw0, ceps, e1, e3 = (float(p_f[k]) for k in (6, 7, 8, 9, 10))
    nc, nmag, depol = int(p_i[3]), int(p_i[10]), int(p_i[13])
    D00, D01, D10, D11 = dmi_matrix(p_i)
    EG = np.zeros_like(F)
    for i in range(nmag):
        for j in range(nc):
            EG[i] = EG[i] - D2[i, j]
    if D00 != 0.0 or D01 != 0.0 or D10 != 0.0 or D11 != 0.0:
        for i in range(nmag):
            for j in range(nmag):
                eps = levi(0, i, j)
                if eps != 0.0:
                    EG[i] = EG[i] + 2.0 * D00 * eps * D1[j, 0]
                if eps != 0.0 and D01 != 0.0:
                    EG[i] = EG[i] + 2.0 * D01 * eps * D1[j, 1]
                eps = levi(1, i, j)
                if eps != 0.0 and D10 != 0.0:
                    EG[i] = EG[i] + 2.0 * D10 * eps * D1[j, 0]
                if eps != 0.0 and D11 != 0.0:
                    EG[i] = EG[i] + 2.0 * D11 * eps * D1[j, 1]
    last = nmag - 1
    EG[last] = EG[last] - PotE * F[last]
    EG[last] = EG[last] - Potw0 * F[last]
    if depol:
        r = e3 / e1
        for j in range(nc):
            EG[3] = EG[3] - D2[3, j]
            for i in range(nc):
                EG[3] = EG[3] + (1.0 / ceps) * (D1[i, i] * D1[j, j] + F[i] * _d2(D2, j, i, j)
                                                - r * (D1[j, i] * D1[i, j] + F[j] * _d2(D2, i, i, j)))
                EG[i] = EG[i] + D1[3, i] * D1[j, j] - D1[3, j] * D1[j, i] - F[j] * _d2(D2, 3, i, j)
                EG[i] = EG[i] + r * (D1[3, i] * D1[j, j] - D1[3, j] * D1[j, i] + F[j] * _d2(D2, 3, i, j))
            for i in range(nmag):
                EG[i] = EG[i] - r * F[i] * D2[3, j]
    EG[:, core.band_mask(p_i)] = 0.0
    core.tangent_project(EG, F, nmag)
    return EG


def normalize(F, p_i, p_f):
    core.unit_norm(F, int(p_i[10]))


def project(V, F, p_i, p_f):
    core.tangent_project(V, F, int(p_i[10]))


def skyrmion_density(F, D1, p_i, p_f):
    """kernels.py:254-287."""
    gv = float(p_f[4])
    cx0 = D1[1, 0] * D1[2, 1] - D1[2, 0] * D1[1, 1]
    cx1 = D1[2, 0] * D1[0, 1] - D1[0, 0] * D1[2, 1]
    cx2 = D1[0, 0] * D1[1, 1] - D1[1, 0] * D1[0, 1]
    return (F[0] * cx0 + F[1] * cx1 + F[2] * cx2) * (gv / (4.0 * math.pi))


def electric_charge_density(F, D1, D2, p_i, p_f):
    """kernels.py:325-362: -(d_i n_i d_j n_j + n_i d_i d_j n_j - (e3/e1)(d_i n_j d_j n_i + n_j d_i d_j n_i)) gv."""
    gv, e1, e3 = float(p_f[4]), float(p_f[9]), float(p_f[10])
    r = e3 / e1
    c = 0.0 * F[0]
    for i in range(2):
        for j in range(2):
            c = c - (D1[i, i] * D1[j, j] + F[i] * _d2(D2, j, i, j) - r * (D1[j, i] * D1[i, j] + F[j] * _d2(D2, i, i, j)))
    return c * gv


CHARGES = {"skyrmion_number": skyrmion_density}
