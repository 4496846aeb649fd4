"""Initial-condition generators of the reference (oracle, FP64) -- TEST
INFRASTRUCTURE ONLY.

Restates, site by site in plain Python/NumPy, the formulas of the
reference's initial-configuration kernels, so that the product's host
generators (``soliton_solver_amd/theories/*/initial_config.py``) and its
device generator (``ss_init_field``) are checked against something other than
themselves:

* chiral magnet and baby Skyrme (identical formulas; skN is p_f[9] resp.
  p_f[8]): ``position`` / ``profile_function_magnetization`` / ``angle``
  (chiral_magnet/initial_config.py:16-115, baby_skyrme/initial_config.py:
  18-117), the centred ansatz of ``create_initial_configuration_kernel``
  (chiral_magnet/initial_config.py:183-284, baby_skyrme/initial_config.py:
  184-281) and the stereographic composition of ``create_skyrmion_kernel``
  (chiral_magnet/initial_config.py:286-386, baby_skyrme/initial_config.py:
  283-383);
* Ginzburg-Landau: ``profile_function_superconductor`` / ``angle``, the
  ground state, the centred vortex of ``create_initial_configuration_kernel``
  and the phase composition of ``create_vortex_kernel``
  (ginzburg_landau_superconductor/initial_config.py:14-313).

The reference normalises m with compute_norm_magnetization
(chiral_magnet/kernels.py:47-80): m /= |m|.
"""
from __future__ import annotations

import math

import numpy as np

EPS = 1e-12


def grid_coords(p_i, p_f):
    """grid[0] = lsx x, grid[1] = lsy y (chiral_magnet/kernels.py:17-45)."""
    X, Y = int(p_i[0]), int(p_i[1])
    g = np.empty((2, X, Y))
    g[0] = float(p_f[2]) * np.arange(X)[:, None]
    g[1] = float(p_f[3]) * np.arange(Y)[None, :]
    return g


def profile_magnetization(r, max_r):
    """pi exp(-(2 r / max_r)^2) for r <= max_r, else 0 (chiral_magnet/initial_config.py:49-72)."""
    if r > max_r:
        return 0.0
    t = 2.0 * r / max_r
    return math.pi * math.exp(-(t * t))


def angle(gx, gy, r, rotation):
    """atan2 of the offset, shifted into [0, 2 pi), minus the rotation; just
    -rotation at the centre (chiral_magnet/initial_config.py:75-115)."""
    theta = -rotation
    if r == 0.0:
        return theta
    a = math.atan2(gy, gx)
    if a < 0.0:
        a += 2.0 * math.pi
    return theta + a


def ansatz_components(kind, fm, th, skN):
    """Bloch / Neel / anti unit vector (chiral_magnet/initial_config.py:249-266,361-376)."""
    if kind == "bloch":
        return -math.sin(fm) * math.sin(skN * th), math.sin(fm) * math.cos(skN * th), math.cos(fm)
    if kind == "neel":
        return math.sin(fm) * math.cos(skN * th), math.sin(fm) * math.sin(skN * th), math.cos(fm)
    if kind == "anti":
        return -math.sin(fm) * math.sin(skN * th), -math.sin(fm) * math.cos(skN * th), math.cos(fm)
    return 0.0, 0.0, 1.0


def _normalise(m1, m2, m3):
    s = math.sqrt(m1 * m1 + m2 * m2 + m3 * m3)
    return m1 / s, m2 / s, m3 / s


def centred_ansatz(p_i, p_f, kind, skN, rotation=0.0):
    """m of create_initial_configuration_kernel: one ansatz at the lattice
    centre, profile radius max(xsize, ysize)/10 * skN; shape (3, X, Y)."""
    X, Y = int(p_i[0]), int(p_i[1])
    g = grid_coords(p_i, p_f)
    xsize, ysize = float(p_f[0]), float(p_f[1])
    xc = (g[0, 0, 0] + g[0, X - 1, Y - 1]) / 2.0
    yc = (g[1, 0, 0] + g[1, X - 1, Y - 1]) / 2.0
    rmax = xsize if xsize > ysize else ysize
    out = np.empty((3, X, Y))
    for x in range(X):
        for y in range(Y):
            gx, gy = g[0, x, y] - xc, g[1, x, y] - yc
            r = math.sqrt(gx * gx + gy * gy)
            th = angle(gx, gy, r, rotation)
            fm = profile_magnetization(r, rmax / 10.0 * skN)
            out[:, x, y] = _normalise(*ansatz_components(kind, fm, th, skN))
    return out


def compose_skyrmion(m, p_i, p_f, xc, yc, max_r, kind, skN, rotation=0.0):
    """Add one soliton to m in place by stereographic addition
    (create_skyrmion_kernel, chiral_magnet/initial_config.py:286-386): w = (m1 + i m2)/(1 + m3)
    with the denominators clamped at 1e-12, w_new = w_old + w_ansatz,
    m = (2 Re w, 2 Im w, 1 - |w|^2)/(1 + |w|^2), then m /= |m|.  skN == 0
    uses a charge-1 ansatz with the doubled profile angle."""
    X, Y = int(p_i[0]), int(p_i[1])
    g = grid_coords(p_i, p_f)
    for x in range(X):
        for y in range(Y):
            gx, gy = g[0, x, y] - xc, g[1, x, y] - yc
            r = math.sqrt(gx * gx + gy * gy)
            th = angle(gx, gy, r, rotation)
            if skN == 0:
                a1, a2, a3 = ansatz_components(kind, 2.0 * profile_magnetization(r, max_r), th, 1.0)
            else:
                a1, a2, a3 = ansatz_components(kind, profile_magnetization(r, max_r), th, skN)
            den = max(1.0 + m[2, x, y], EPS)
            wr, wi = m[0, x, y] / den, m[1, x, y] / den
            iden = max(1.0 + a3, EPS)
            wr, wi = wr + a1 / iden, wi + a2 / iden
            nw = wr * wr + wi * wi
            m[:, x, y] = _normalise(2.0 * wr / (1.0 + nw), 2.0 * wi / (1.0 + nw), (1.0 - nw) / (1.0 + nw))
    return m


def profile_superconductor(r, max_r, value):
    """tanh(value r) for r <= max_r, else 1 (ginzburg_landau_superconductor/initial_config.py:47-72)."""
    if r > max_r:
        return 1.0
    return math.tanh(value * r)


def gl_ground(p_i, p_f):
    """A = 0, psi = u1 (create_ground_state_kernel, :115-179; u1 = p_f[8])."""
    X, Y = int(p_i[0]), int(p_i[1])
    F = np.zeros((4, X, Y))
    F[2] = float(p_f[8])
    return F


def gl_centred_vortex(p_i, p_f, vortex_number, ainf):
    """create_initial_configuration_kernel of the Ginzburg-Landau theory (:181-258): at
    the lattice centre, fs = tanh(0.2 r) (1 beyond max(xsize, ysize)),
    A = ainf fs (sin th, -cos th)/max(r, 1e-12), psi = u1 fs e^{i N th}.
    ``ainf`` is the packed A-infinity value (the reference reads it one past
    the packed array, p_f[11]; the product reads p_f[10])."""
    X, Y = int(p_i[0]), int(p_i[1])
    g = grid_coords(p_i, p_f)
    xsize, ysize, u1 = float(p_f[0]), float(p_f[1]), float(p_f[8])
    xc = (g[0, 0, 0] + g[0, X - 1, Y - 1]) / 2.0
    yc = (g[1, 0, 0] + g[1, X - 1, Y - 1]) / 2.0
    rmax = xsize if xsize > ysize else ysize
    F = np.empty((4, X, Y))
    for x in range(X):
        for y in range(Y):
            gx, gy = g[0, x, y] - xc, g[1, x, y] - yc
            r = math.sqrt(gx * gx + gy * gy)
            th = angle(gx, gy, r, 0.0)
            fs = profile_superconductor(r, rmax, 0.2)
            rr = r if r > EPS else EPS
            F[0, x, y] = ainf * fs * math.sin(th) / rr
            F[1, x, y] = -ainf * fs * math.cos(th) / rr
            F[2, x, y] = u1 * fs * math.cos(vortex_number * th)
            F[3, x, y] = u1 * fs * math.sin(vortex_number * th)
    return F


def gl_compose_vortex(F, p_i, p_f, xc, yc, antivortex=False):
    """create_vortex_kernel (:260-313): psi <- psi * u1 fs e^{i N phase} / u1, with
    phase = -th for an antivortex, N = p_f[9]; A unchanged."""
    X, Y = int(p_i[0]), int(p_i[1])
    g = grid_coords(p_i, p_f)
    xsize, ysize, u1, N = float(p_f[0]), float(p_f[1]), float(p_f[8]), float(p_f[9])
    rmax = xsize if xsize > ysize else ysize
    for x in range(X):
        for y in range(Y):
            gx, gy = g[0, x, y] - xc, g[1, x, y] - yc
            r = math.sqrt(gx * gx + gy * gy)
            th = angle(gx, gy, r, 0.0)
            fs = profile_superconductor(r, rmax, 0.2)
            ph = (-1.0 if antivortex else 1.0) * th
            ir, ii = u1 * fs * math.cos(N * ph), u1 * fs * math.sin(N * ph)
            orr, oi = F[2, x, y], F[3, x, y]
            F[2, x, y] = (orr * ir - oi * ii) / u1
            F[3, x, y] = (orr * ii + oi * ir) / u1
    return F
