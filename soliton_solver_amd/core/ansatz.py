"""Host-side seeded initial conditions for the unit-vector theories.

The formulas are the reference's initial-condition kernels, evaluated with
NumPy on the host (the relaxation core receives the resulting arrays):

* radial profile ``pi * exp(-(2 r / max_r)^2)`` for r <= max_r, else 0
  (chiral_magnet/initial_config.py:48-72);
* polar angle ``atan2`` shifted to [0, 2 pi) minus the rotation, and just
  ``-rotation`` at r == 0 (initial_config.py:74-115);
* Bloch / Neel / anti ansatz components (initial_config.py:249-266,361-376);
* placement by adding stereographic coordinates w = (m1 + i m2)/(1 + m3)
  (denominators clamped at 1e-12) and mapping back, then unit-normalising
  (initial_config.py:285-386).

Array layout is ``F[a, x, y]`` with ``grid[0] = lsx*x``, ``grid[1] = lsy*y``.
"""
from __future__ import annotations

import numpy as np

_EPS = 1e-12


def profile(r, max_r):
    t = 2.0 * r / max_r
    return np.where(r > max_r, 0.0, np.pi * np.exp(-(t * t)))


def polar_angle(grid, xcent, ycent, rotation):
    gx = grid[0] - xcent
    gy = grid[1] - ycent
    r = np.sqrt(gx * gx + gy * gy)
    ang = np.arctan2(gy, gx)
    ang = np.where(ang < 0.0, ang + 2.0 * np.pi, ang)
    theta = -rotation + ang
    return r, np.where(r == 0.0, -rotation, theta)


def ansatz_vector(fm, th, skN, kind):
    """Unit vector of a Bloch, Neel or anti ansatz with profile angle fm."""
    if kind == "bloch":
        return -np.sin(fm) * np.sin(skN * th), np.sin(fm) * np.cos(skN * th), np.cos(fm)
    if kind == "neel":
        return np.sin(fm) * np.cos(skN * th), np.sin(fm) * np.sin(skN * th), np.cos(fm)
    if kind == "anti":
        return -np.sin(fm) * np.sin(skN * th), -np.sin(fm) * np.cos(skN * th), np.cos(fm)
    z = np.zeros_like(fm)
    return z, z, np.ones_like(fm)


def unit_normalize(F, nmag=3):
    s = np.sqrt(sum(F[a] * F[a] for a in range(nmag)))
    nz = s != 0.0
    for a in range(nmag):
        F[a] = np.where(nz, F[a] / np.where(nz, s, 1.0), F[a])


def single_ansatz(F, grid, xcent, ycent, max_r, rotation, skN, kind):
    """Overwrite m with one ansatz centred at (xcent, ycent) (initial_config.py:182-266)."""
    r, th = polar_angle(grid, xcent, ycent, rotation)
    fm = profile(r, max_r)
    m1, m2, m3 = ansatz_vector(fm, th, skN, kind)
    F[0], F[1], F[2] = m1, m2, m3
    unit_normalize(F, 3)


def place_soliton(F, grid, xcent, ycent, max_r, rotation, skN, kind, window=True):
    """Compose one more soliton into m by stereographic addition (initial_config.py:285-386).

    The profile vanishes for r > max_r, where the composition is the identity
    (up to the round trip through stereographic coordinates); with ``window``
    only the bounding box of the disc (plus 2 cells) is evaluated.
    """
    if window:
        X, Y = F.shape[1], F.shape[2]
        lsx = grid[0, 1, 0] - grid[0, 0, 0] if X > 1 else 1.0
        lsy = grid[1, 0, 1] - grid[1, 0, 0] if Y > 1 else 1.0
        x0 = max(0, int(np.floor((xcent - max_r) / lsx)) - 2)
        x1 = min(X, int(np.ceil((xcent + max_r) / lsx)) + 3)
        y0 = max(0, int(np.floor((ycent - max_r) / lsy)) - 2)
        y1 = min(Y, int(np.ceil((ycent + max_r) / lsy)) + 3)
        if x0 >= x1 or y0 >= y1:
            return
        sub = F[:, x0:x1, y0:y1].copy()
        place_soliton(sub, grid[:, x0:x1, y0:y1], xcent, ycent, max_r, rotation, skN, kind, window=False)
        F[:, x0:x1, y0:y1] = sub
        return
    r, th = polar_angle(grid, xcent, ycent, rotation)
    if skN == 0:
        fm = 2.0 * profile(r, max_r)
        m1, m2, m3 = ansatz_vector(fm, th, 1.0, kind)
    else:
        fm = profile(r, max_r)
        m1, m2, m3 = ansatz_vector(fm, th, skN, kind)
    den = np.maximum(1.0 + F[2], _EPS)
    old_wr, old_wi = F[0] / den, F[1] / den
    int_den = np.maximum(1.0 + m3, _EPS)
    wr = old_wr + m1 / int_den
    wi = old_wi + m2 / int_den
    nw = wr * wr + wi * wi
    F[0] = 2.0 * wr / (1.0 + nw)
    F[1] = 2.0 * wi / (1.0 + nw)
    F[2] = (1.0 - nw) / (1.0 + nw)
    unit_normalize(F, 3)


def gaussian_perturb(F, rng, sigma, nmag=3):
    """m += N(0, sigma^2) per component, then renormalise |m| = 1."""
    for a in range(nmag):
        F[a] = F[a] + rng.normal(0.0, sigma, size=F[a].shape)
    unit_normalize(F, nmag)
