"""Golden vectors of the 320 x 320 chiral-magnet relaxation (the drop-in FP64
workload), made with the oracle's C++/OpenMP restatement of the reference
path (oracle/cpu_ref.cpp; core/integrator.py:217-348 driven as
chiral_magnet/render_gl.py:647-670: the first prev_energy is the energy of
the initial field).

Workload: 320 x 320, xsize = ysize = 10, J, K, D, M, B = 40e-12, 0.8e6,
4e-3, 580e3, 0, Heusler DMI, demagnetisation on, unit magnetisation; the
reference's centred anti-skyrmion ansatz (chiral_magnet/initial_config.py:
183-284) plus a seeded Gaussian perturbation sigma = 1e-3 on m (NumPy
default_rng(seed).normal, components 1..3 in order), renormalised; psi = 0;
run for 150 000 arrested-Newton steps.  (The flow crosses max |stage-4
energy gradient| = 1e-6 near step 80 000 and then sits on a plateau: from step
140 000 the energy falls by ~4e-13 per 1000 steps while max |EG| stays at
7.2-7.3e-7 -- the noise-displaced anti-skyrmion slides along a nearly flat
direction -- so 1e-8 is out of reach of the reference algorithm here; the
fixture records the state after a fixed step count, which a faithful
implementation reproduces to the plateau's energy.)

Run from the repository root: ``python tests/golden/make_chiral320.py``
(about 8 minutes on 12 cores of an EPYC host).  Writes tests/golden/chiral320_relaxed.json.
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

N, XSIZE, SEED, SIGMA, NSTEPS = 320, 10.0, 11, 1e-3, 150000
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "chiral320_relaxed.json")
SAMPLE = slice(8, N, 16)  # 20 x 20 sample sites


def params():
    import soliton_solver_amd as ssa
    th = ssa.load_theory("Chiral magnet")
    p = th.params.default_params(xlen=N, ylen=N, xsize=XSIZE, ysize=XSIZE, J=40e-12, K=0.8e6, D=4e-3, M=580e3,
                                 B=0.0, dmi_term="Heusler", demag=True, unit_magnetization=True)
    p_i, p_f = th.params.pack_device_params(p.resolved())
    return th, p, np.ascontiguousarray(p_i, dtype=np.int32), np.ascontiguousarray(p_f, dtype=np.float64)


def initial_field(p_i, p_f):
    from oracle import initial as oinit
    F = np.zeros((4, N, N))
    F[:3] = oinit.centred_ansatz(p_i, p_f, "anti", float(p_f[9]))
    rng = np.random.default_rng(SEED)
    for a in range(3):
        F[a] = F[a] + rng.normal(0.0, SIGMA, size=(N, N))
    F[:3] /= np.sqrt((F[:3] ** 2).sum(0))
    return F


def summary(F, p_i, p_f):
    from oracle import core as ocore
    from oracle.theories import chiral_magnet as o_cm
    D1 = ocore.first_derivatives(F, p_i, p_f)
    Q = float(np.sum(o_cm.skyrmion_density(F, D1, p_i, p_f)))
    return {"skyrmion_number": Q, "energy_oracle": ocore.theory_energy(o_cm, F, p_i, p_f),
            "sum_m": [float(F[a].sum()) for a in range(3)], "sum_psi": float(F[3].sum()),
            "sample_m": F[:3, SAMPLE, SAMPLE].round(12).tolist(), "sample_psi": F[3, SAMPLE, SAMPLE].round(12).tolist()}


def main():
    from oracle import core as ocore
    from oracle import cpu_build
    from oracle.theories import chiral_magnet as o_cm
    th, p, p_i, p_f = params()
    F = np.ascontiguousarray(initial_field(p_i, p_f))
    V = np.zeros_like(F)
    E0 = ocore.theory_energy(o_cm, F, p_i, p_f)
    init = {"energy": E0, "sum_m": [float(F[a].sum()) for a in range(3)]}
    t0 = time.perf_counter()
    E, steps, ea = E0, 0, np.inf
    trace = []
    while steps < NSTEPS:
        Es, errs, eas = cpu_build.chiral_relax_to(p_i, p_f, F, V, E, min(20000, NSTEPS - steps), 0.0)
        steps += len(Es)
        E, ea = float(Es[-1]), float(eas[-1])
        trace.append([steps, E, ea])
        print(f"{steps} steps, {time.perf_counter() - t0:.0f} s: E={E:.15f} max|EG|={ea:.3e}", flush=True)
    out = {"workload": __doc__.split("\n\n")[1].replace("\n", " "), "seed": SEED, "sigma": SIGMA, "trace": trace,
           "lattice": N, "initial": init, "steps": steps, "energy": E, "err_absmax": ea,
           "sample_rows_cols": [SAMPLE.start, SAMPLE.stop, SAMPLE.step], **summary(F, p_i, p_f),
           "seconds": time.perf_counter() - t0}
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {OUT}: E={E:.15f}, Q={out['skyrmion_number']:.9f}, {steps} steps", flush=True)


if __name__ == "__main__":
    main()
