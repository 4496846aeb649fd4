"""Auxiliary per-site quantities (oracle, FP64).

TEST INFRASTRUCTURE ONLY.  Restates the reference's compute_norm_higgs*_point,
compute_phase_difference_point and compute_supercurrent_point of the gauge
theories (e.g. ginzburg_landau_superconductor/kernels.py,
anisotropic_superconductor/kernels.py, ferromagnetic_superconductor/kernels.py)
and their Simulation wrappers (core/simulation.py:211-386).  D2[a] holds
(d_xx, d_yy, mixed) of component a, the mixed entry shared by (x,y) and (y,x).
"""
from __future__ import annotations

import numpy as np


def higgs(F, re, im):
    """|psi|^2 of the order parameter stored at components re, im."""
    return F[re] * F[re] + F[im] * F[im]


def phase_difference(F, p1=(2, 3), p2=(4, 5)):
    """cos(atan2(psi1_im, psi1_re) - atan2(psi2_im, psi2_re)) (anisotropic_superconductor)."""
    return np.cos(np.arctan2(F[p1[1]], F[p1[0]]) - np.arctan2(F[p2[1]], F[p2[0]]))


def supercurrent(D2, a1, a2, a3=None):
    """(d_yx A2 - d_yy A1, d_xy A1 - d_xx A2[, -(d_xx + d_yy) A3]) for gauge components a1, a2, a3."""
    out = [D2[a2, 2] - D2[a1, 1], D2[a1, 2] - D2[a2, 0]]
    if a3 is not None:
        out.append(-(D2[a3, 0] + D2[a3, 1]))
    return np.stack(out)
