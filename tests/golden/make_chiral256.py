"""Golden vectors of a 256 x 256 chiral-magnet relaxation at the spacing of
the 2048^2 / 4096^2 precision workloads, made with the oracle's C++/OpenMP
FP64 restatement of the reference path (oracle/cpu_ref.cpp;
core/integrator.py:217-348 driven as chiral_magnet/render_gl.py:647-670).

Workload: 256 x 256 at h = 10/319 (xsize = ysize = 255 h), J, K, D, M, B =
40e-12, 0.8e6, 4e-3, 580e3, 0, Heusler DMI, demagnetisation off, unit
magnetisation; the ground state with one anti-skyrmion of radius R = 0.45 at
the lattice centre composed stereographically (chiral_magnet/initial_config.py:
286-386, profile radius 2R) plus a seeded Gaussian perturbation sigma = 1e-3
on m (NumPy default_rng(seed).normal, components 1..3 in order),
renormalised; psi = 0; run for 200 000 arrested-Newton steps (max |stage-4
energy gradient| ~1.4e-6 there, decreasing by ~5 % per 20 000 steps; the
energy changes by ~5e-10 per 20 000 steps, 6e-11 relative).

Run from the repository root: ``python tests/golden/make_chiral256.py``.
Writes tests/golden/chiral256_relaxed.json.
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N, SEED, SIGMA, NSTEPS, RADIUS = 256, 21, 1e-3, 200000, 0.45
H_SPACING = 10.0 / 319.0
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "chiral256_relaxed.json")
SAMPLE = slice(8, N, 16)  # 16 x 16 sample sites


def params():
    import soliton_solver_amd as ssa
    th = ssa.load_theory("Chiral magnet")
    L = H_SPACING * (N - 1)
    p = th.params.default_params(xlen=N, ylen=N, xsize=L, ysize=L, J=40e-12, K=0.8e6, D=4e-3, M=580e3, B=0.0,
                                 dmi_term="Heusler", demag=False, unit_magnetization=True)
    p_i, p_f = th.params.pack_device_params(p.resolved())
    return th, p, np.ascontiguousarray(p_i, dtype=np.int32), np.ascontiguousarray(p_f, dtype=np.float64)


def initial_field(p_i, p_f):
    from oracle import initial as oinit
    L = H_SPACING * (N - 1)
    m = np.zeros((3, N, N))
    m[2] = 1.0
    oinit.compose_skyrmion(m, p_i, p_f, L / 2.0, L / 2.0, 2.0 * RADIUS, "anti", float(p_f[9]))
    rng = np.random.default_rng(SEED)
    for a in range(3):
        m[a] = m[a] + rng.normal(0.0, SIGMA, size=(N, N))
    m /= np.sqrt((m ** 2).sum(0))
    F = np.zeros((4, N, N))
    F[:3] = m
    return F


def main():
    from oracle import core as ocore
    from oracle import cpu_build
    from oracle.theories import chiral_magnet as o_cm
    th, p, p_i, p_f = params()
    F = np.ascontiguousarray(initial_field(p_i, p_f))
    V = np.zeros_like(F)
    E0 = ocore.theory_energy(o_cm, F, p_i, p_f)
    t0 = time.perf_counter()
    E, steps, ea = E0, 0, np.inf
    trace = []
    while steps < NSTEPS:
        Es, errs, eas = cpu_build.chiral_relax_to(p_i, p_f, F, V, E, min(20000, NSTEPS - steps), 0.0)
        steps += len(Es)
        E, ea = float(Es[-1]), float(eas[-1])
        trace.append([steps, E, ea])
        print(f"{steps} steps, {time.perf_counter() - t0:.0f} s: E={E:.15f} max|EG|={ea:.3e}", flush=True)
    D1 = ocore.first_derivatives(F, p_i, p_f)
    out = {"workload": __doc__.split("\n\n")[1].replace("\n", " "), "seed": SEED, "sigma": SIGMA, "trace": trace,
           "lattice": N, "radius": RADIUS, "initial_energy": E0, "steps": steps, "energy": E, "err_absmax": ea,
           "skyrmion_number": float(np.sum(o_cm.skyrmion_density(F, D1, p_i, p_f))),
           "sample_rows_cols": [SAMPLE.start, SAMPLE.stop, SAMPLE.step],
           "sample_m": F[:3, SAMPLE, SAMPLE].round(12).tolist(), "seconds": time.perf_counter() - t0}
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {OUT}: E={E:.15f}, Q={out['skyrmion_number']:.9f}, {steps} steps", flush=True)


if __name__ == "__main__":
    main()
