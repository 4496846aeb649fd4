"""Core numerics of the oracle: stencils, reductions, RK4 arrested Newton flow.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Every function restates one reference function; the citation is on each.
Arithmetic is FP64 and follows the reference's left-to-right evaluation order
so that per-stage results agree with a faithful FP64 implementation to the
last few ulps.
"""
from __future__ import annotations

import numpy as np

C12 = 1.0 / 12.0
C23 = 2.0 / 3.0
C43 = 4.0 / 3.0
C52 = 5.0 / 2.0


def _dims(p_i):
    X, Y, halo = int(p_i[0]), int(p_i[1]), int(p_i[2])
    if halo < 2:
        raise ValueError("halo must be >= 2 for the 5-point stencils")
    return X, Y, halo


def first_derivatives(F, p_i, p_f):
    """4th-order central first derivatives, zero inside the halo band.

    Restates core/derivatives.py:13-57: d/dx is
    ``((1/12)f[x-2] - (2/3)f[x-1] + (2/3)f[x+1] - (1/12)f[x+2]) / lsx`` for
    ``halo-1 < x < xlen-halo`` and 0 elsewhere; likewise d/dy with lsy.
    Returns ``D1[a, i, x, y]``.
    """
    X, Y, h = _dims(p_i)
    lsx, lsy = float(p_f[2]), float(p_f[3])
    nf = F.shape[0]
    D1 = F.new_zeros((nf, 2, X, Y)) if hasattr(F, "new_zeros") else np.zeros((nf, 2, X, Y))
    xs = slice(h, X - h)
    ys = slice(h, Y - h)
    if X - h > h:
        fm2 = F[:, h - 2:X - h - 2, :]
        fm1 = F[:, h - 1:X - h - 1, :]
        fp1 = F[:, h + 1:X - h + 1, :]
        fp2 = F[:, h + 2:X - h + 2, :]
        D1[:, 0, xs, :] = (C12 * fm2 - C23 * fm1 + C23 * fp1 - C12 * fp2) / lsx
    if Y - h > h:
        fm2 = F[:, :, h - 2:Y - h - 2]
        fm1 = F[:, :, h - 1:Y - h - 1]
        fp1 = F[:, :, h + 1:Y - h + 1]
        fp2 = F[:, :, h + 2:Y - h + 2]
        D1[:, 1, :, ys] = (C12 * fm2 - C23 * fm1 + C23 * fp1 - C12 * fp2) / lsy
    return D1


def second_derivatives(F, p_i, p_f):
    """4th-order second derivatives with the reference's masks and mixed form.

    Restates core/derivatives.py:59-115:
    * xx: ``(-(1/12)f[x-2] + (4/3)f[x-1] - (5/2)f[x] + (4/3)f[x+1] - (1/12)f[x+2])
      / (lsx*lsx)`` for ``halo-1 < x < xlen-halo``, else 0 (yy likewise);
    * mixed: ``diag/(2 gv) - xx_raw/(2 lsx lsx) - yy_raw/(2 lsy lsy)`` where
      diag is the same 5-point second difference along (1,1); only where
      ``x > halo`` and ``y > halo`` (strict) and ``x < xlen-halo``,
      ``y < ylen-halo``; else 0.  The reference stores the same value in
      both off-diagonal slots (derivatives.py:115).
    Returns ``D2[a, k, x, y]`` with k = 0 xx, 1 yy, 2 mixed.
    """
    X, Y, h = _dims(p_i)
    lsx, lsy, gv = float(p_f[2]), float(p_f[3]), float(p_f[4])
    nf = F.shape[0]
    D2 = np.zeros((nf, 3, X, Y))
    if X - h > h:
        sx = slice(h, X - h)
        f0 = F[:, sx, :]
        fm2 = F[:, h - 2:X - h - 2, :]
        fm1 = F[:, h - 1:X - h - 1, :]
        fp1 = F[:, h + 1:X - h + 1, :]
        fp2 = F[:, h + 2:X - h + 2, :]
        D2[:, 0, sx, :] = (-C12 * fm2 + C43 * fm1 - C52 * f0 + C43 * fp1 - C12 * fp2) / (lsx * lsx)
    if Y - h > h:
        sy = slice(h, Y - h)
        f0 = F[:, :, sy]
        fm2 = F[:, :, h - 2:Y - h - 2]
        fm1 = F[:, :, h - 1:Y - h - 1]
        fp1 = F[:, :, h + 1:Y - h + 1]
        fp2 = F[:, :, h + 2:Y - h + 2]
        D2[:, 1, :, sy] = (-C12 * fm2 + C43 * fm1 - C52 * f0 + C43 * fp1 - C12 * fp2) / (lsy * lsy)
    lo = h + 1
    if X - h > lo and Y - h > lo:
        def S(dx, dy):
            return F[:, lo + dx:X - h + dx, lo + dy:Y - h + dy]
        diag = (-C12 * S(-2, -2) + C43 * S(-1, -1) - C52 * S(0, 0) + C43 * S(1, 1) - C12 * S(2, 2)) / (2.0 * gv)
        dxx_half = (-C12 * S(-2, 0) + C43 * S(-1, 0) - C52 * S(0, 0) + C43 * S(1, 0) - C12 * S(2, 0)) / (2.0 * lsx * lsx)
        dyy_half = (-C12 * S(0, -2) + C43 * S(0, -1) - C52 * S(0, 0) + C43 * S(0, 1) - C12 * S(0, 2)) / (2.0 * lsy * lsy)
        D2[:, 2, lo:X - h, lo:Y - h] = diag - dxx_half - dyy_half
    return D2


def reference_sum(values):
    """Sum in the reference's order: 1024-wide shared-memory trees, then numpy.

    Restates core/utils.py:187-228 (tree with halving stride inside each
    1024-thread block, zero padding past ``size``) and utils.py:316-341
    (``partial.copy_to_host().sum()``).
    """
    v = np.asarray(values, dtype=np.float64).ravel()
    n = v.size
    nb = max(1, (n + 1023) // 1024)
    a = np.zeros(nb * 1024)
    a[:n] = v
    a = a.reshape(nb, 1024)
    stride = 512
    while stride > 0:
        a[:, :stride] += a[:, stride:2 * stride]
        stride //= 2
    return float(a[:, 0].sum())


def reference_max(values):
    """Signed max in the reference's order (utils.py:230-271,343-368,397-420).

    The block tree keeps ``sdata[tid+stride]`` only when strictly larger and
    seeds padding with -1e300; for finite inputs that is ``max``.  NaNs are
    not propagated by the tree's ``<`` test, so they are dropped here too.
    """
    v = np.asarray(values, dtype=np.float64).ravel()
    finite = v[~np.isnan(v)]
    if finite.size == 0:
        return -1.0e300
    return float(max(finite.max(), -1.0e300))


def theory_energy(th, F, p_i, p_f):
    """Total energy: per-site density summed in the reference order.

    Restates e.g. chiral_magnet/observables.py:15-57 (density kernel, then
    ``compute_sum`` over ``dim_grid`` entries in ``idx_field(0, x, y)`` order).
    """
    D1 = first_derivatives(F, p_i, p_f)
    e = th.energy_density(F, D1, p_i, p_f)
    return reference_sum(e.ravel())


def gradient_stage(th, Temp, p_i, p_f):
    """One gradient kernel: returns ``(EG, l)`` with ``l = -dt * EG``.

    Restates core/integrator.py:15-48 with the theory's
    ``do_gradient_step_point`` (e.g. chiral_magnet/kernels.py:397-511).
    """
    D1 = first_derivatives(Temp, p_i, p_f)
    D2 = second_derivatives(Temp, p_i, p_f)
    EG = th.variation(Temp, D1, D2, p_i, p_f)
    dt = float(p_f[5])
    return EG, -dt * EG


def arrested_newton_flow_step(th, F, V, p_i, p_f, prev_energy):
    """One arrested-Newton RK4 step, in place on ``F`` and ``V``.

    Restates core/integrator.py:217-348 (the stage kernels :50-96, the
    combine :98-149/:175-187, the optional L2 normalisation :295-321, energy,
    velocity kill with strict ``>`` against the caller's ``prev_energy``,
    and the signed max of the stage-4 energy gradient).
    Returns ``(new_energy, err)``.
    """
    dt = float(p_f[5])
    norm = getattr(th, "l2_normalize", None)

    def maybe_normalize(G):
        if norm is not None:
            norm(G, p_i, p_f)

    k1 = dt * (V + 0.0 * V)
    Temp = F + 0.0 * F
    maybe_normalize(Temp)
    _, l1 = gradient_stage(th, Temp, p_i, p_f)
    k2 = dt * (V + 0.5 * l1)
    Temp = F + 0.5 * k1
    maybe_normalize(Temp)
    _, l2 = gradient_stage(th, Temp, p_i, p_f)
    k3 = dt * (V + 0.5 * l2)
    Temp = F + 0.5 * k2
    maybe_normalize(Temp)
    _, l3 = gradient_stage(th, Temp, p_i, p_f)
    k4 = dt * (V + 1.0 * l3)
    Temp = F + 1.0 * k3
    maybe_normalize(Temp)
    EG, l4 = gradient_stage(th, Temp, p_i, p_f)
    V += (1.0 / 6.0) * (l1 + 2.0 * l2 + 2.0 * l3 + l4)
    F += (1.0 / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    if int(p_i[9]) and getattr(th, "HAS_CONSTRAINT", False):
        th.normalize(F, p_i, p_f)
        th.project(V, F, p_i, p_f)
    maybe_normalize(F)
    new_energy = theory_energy(th, F, p_i, p_f)
    if int(p_i[7]) and new_energy > prev_energy:
        V[...] = 0.0
    err = reference_max(EG.ravel())
    return new_energy, err


def relax(th, F, V, p_i, p_f, nsteps, prev_energy=None):
    """Drive ``nsteps`` steps the way the reference's viewer loop does.

    Restates chiral_magnet/render_gl.py:647-670 (``advance_solver``): the
    first ``prev_energy`` is the energy of the initial field, then
    ``E, err = step(E)``.  Returns lists of per-step energies and errors.
    """
    if prev_energy is None:
        prev_energy = theory_energy(th, F, p_i, p_f)
    Es, errs = [], []
    E = prev_energy
    for _ in range(nsteps):
        E, err = arrested_newton_flow_step(th, F, V, p_i, p_f, E)
        Es.append(E)
        errs.append(err)
    return Es, errs


def unit_norm(F, nmag):
    """Pointwise normalisation of the first ``nmag`` components.

    Restates chiral_magnet/kernels.py:47-83 (``compute_norm_magnetization``):
    ``s = sqrt(sum m_a^2)``; sites with ``s == 0`` are left unchanged.
    """
    s = np.zeros(F.shape[1:])
    for a in range(nmag):
        s = s + F[a] * F[a]
    s = np.sqrt(s)
    safe = np.where(s == 0.0, 1.0, s)
    for a in range(nmag):
        F[a] = np.where(s == 0.0, F[a], F[a] / safe)


def tangent_project(G, F, nmag):
    """``G -= (G.m) m`` over the first ``nmag`` components.

    Restates chiral_magnet/kernels.py:85-119 (``project_orthogonal_magnetization``).
    """
    lm = np.zeros(F.shape[1:])
    for a in range(nmag):
        lm = lm + G[a] * F[a]
    for a in range(nmag):
        G[a] = G[a] - lm * F[a]


def band_mask(p_i):
    """Boolean mask of sites in the 2-site boundary band.

    Restates the test at chiral_magnet/kernels.py:505-507:
    ``x < halo or x > xlen-halo-1 or y < halo or y > ylen-halo-1``.
    """
    X, Y, h = _dims(p_i)
    m = np.zeros((X, Y), dtype=bool)
    m[:h, :] = True
    m[X - h:, :] = True
    m[:, :h] = True
    m[:, Y - h:] = True
    return m
