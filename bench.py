"""Benchmark of the relaxation core on one MI355X.

Workload (default): chiral magnet, 4096 x 4096 lattice at fixed spacing
h = 10/319 (xsize = 4095 h), Heusler DMI, demagnetisation on, unit
magnetisation, coupling constants J, K, D, M, B = 40e-12, 0.8e6, 4e-3, 580e3, 0;
initial field = 8 seeded anti-skyrmions (centres uniform in the central 80 %,
radii uniform in [0.3, 0.6]) plus seeded Gaussian noise sigma = 1e-3,
renormalised.  One step = one arrested-Newton RK4 step over every site.

The headline ``value`` is grid-point updates/s of the FP32-stencil /
FP64-reduction form that converges to the FP64 minimum: the compensated form
(FP32 fields and stencils, the field held as FP32 hi + lo words so that
sub-ulp increments accumulate, FP64 energy of hi + lo, FP64 sums and arrest
test).  The plain FP32 form and the plain FP64 form are timed beside it on the
same initial field; the plain FP32 form alone does not reach Delta E / E <=
1e-7 without a long FP64 tail.  ``hbm_frac_alg`` prices every form at the
fixed algorithmic 16 nf w bytes per update (w = 4 for FP32 fields, 8 for
FP64); ``hbm_frac_schedule`` at the words the chosen schedule actually moves.
``roofline`` prices the dominant kernel (the launch that finishes a step:
pass 4, the 4+1 sweep or the 3+4 sweep, at the words it reads and writes)
from HIP events on the context's stream.  ``time_to_converge`` relaxes one
4096^2 anti-skyrmion (demagnetisation off) to max |dE/dphi| < 1e-6 in the
compensated form + FP64 polish and in the plain FP64 form.
``cpu_baseline`` times the oracle's C++/OpenMP FP64 restatement of the
reference path on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "grid-pt updates/s + %HBM peak; FP32/FP64-reduce vs FP64 at ΔE/E≤1e-7"


def chiral_workload(n: int, seed: int = 0, nsk: int = 8, demag: bool = True):
    import soliton_solver_amd as ssa
    th = ssa.load_theory("Chiral magnet")
    h = 10.0 / 319.0
    L = h * (n - 1)
    p = th.params.default_params(xlen=n, ylen=n, xsize=L, ysize=L, J=40e-12, K=0.8e6, D=4e-3, M=580e3, B=0.0,
                                 dmi_term="Heusler", demag=demag, unit_magnetization=True)
    rng = np.random.default_rng(seed)
    sk = [(rng.uniform(0.1 * L, 0.9 * L), rng.uniform(0.1 * L, 0.9 * L), rng.uniform(0.3, 0.6))
          for _ in range(nsk)]
    return th, p, {"mode": "skyrmions", "ansatz": "anti", "skyrmions": sk, "noise": 1e-3, "seed": seed + 1}


def time_form(th, p, cfg, precision, steps, warmup, graph_steps, F0=None, autotune=True, tiles=None, fused23=-1,
              fused41=-1, pairs=-1):
    import soliton_solver_amd as ssa
    sim = ssa.Simulation(p, th, precision=precision)
    if F0 is None:
        sim.initialize(cfg)
    else:
        sim.initialize({"mode": "array", "field": F0})
    if tiles is not None:
        sim.set_tile_variants(tiles)
        if fused23 >= 0 or fused41 >= 0 or pairs >= 0:
            sim.set_fused_stages(fused23 > 0, fused41 > 0, pairs=pairs > 0)
        sim.tiles = sim.tile_variants()
    else:
        sim.tiles = sim.autotune_tiles(reps=20) if autotune else sim.tile_variants()
    if graph_steps:
        sim.use_graphs(graph_steps)
    E = sim.compute_energy()
    E, _, _ = sim.steps(E, warmup)
    # the graphs the timed call replays are captured and uploaded here, not in the timed region
    sim.prepare_graphs(steps)
    sim.synchronize()
    import torch
    torch.cuda.synchronize()
    E_start = E
    t0 = time.perf_counter()
    E, err, ea = sim.steps(E, steps)
    sim.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # how many timed steps ran with the velocity arrested (they read no V)
    sim.timed_arrested = int(arrested_steps(sim.history(steps)[0], E_start).sum()) if steps > 0 else 0
    return sim, E, err, ea, dt


def relax_timed(sim, E, tol, max_steps, check_every, floor=False):
    """Relax until max |EG| < tol (or, with ``floor``, until max |EG| stops
    improving: three checks without a new minimum).  Returns (E, err_abs, steps, seconds)."""
    import torch
    sim.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done, best, stale = 0, np.inf, 0
    ea = np.inf
    while done < max_steps:
        n = min(check_every, max_steps - done)
        E, _, ea = sim.steps(E, n)
        done += n
        if done % (20 * check_every) == 0:
            print(f"[bench]   {done} steps, E={E:.12f}, max|EG|={ea:.2e}", file=sys.stderr, flush=True)
        if ea < tol:
            break
        if floor:
            if ea < 0.98 * best:
                best, stale = ea, 0
            else:
                stale += 1
                if stale >= 3:
                    break
    sim.synchronize()
    return E, ea, done, time.perf_counter() - t0


def precision_workload(args, emit=True):
    """Time to max |EG| < tol: the plain FP64 form against the FP32 form with
    FP64 polish, from the same seeded field.  ``--form fp32c`` (the default)
    runs the FP32 phase with the compensated field update (FP32 hi + lo words,
    increments added in FP64), which keeps the steps whose increments are
    below an FP32 ulp of the field; ``--form fp32`` runs the plain FP32 form.

    FP32 storage puts a floor under max |EG| (field rounding ~6e-8 over h^2
    ~ 1e-3 gives ~1e-4), but the slow modes keep converging below it.  The
    FP32 form therefore runs in chunks of at most ``--chunk`` steps; after each
    chunk the state is converted to FP64 in place and polished for at most
    ``--polish`` steps: if max |EG| < tol it is done, otherwise the state goes
    back to FP32 and the next chunk runs, sized from the decay rate of the
    polished max |EG| over the last two checks.  When a chunk improves the
    polished max |EG| by less than 30 %, FP32 has stalled (its steps no longer
    resolve the remaining increments) and the run finishes in FP64 (the tail,
    counted with the polish time)."""
    import soliton_solver_amd as ssa
    import torch
    n = args.n
    th, p, cfg = chiral_workload(n, seed=args.seed, nsk=args.nsk, demag=bool(args.demag))
    tol = args.tol
    s64 = ssa.Simulation(p, th, precision="fp64")
    s64.initialize(cfg)
    F0 = s64.get_field()
    s64.autotune_tiles(reps=3)
    s64.use_graphs(args.graph_steps)
    E0 = s64.compute_energy()
    E64, ea64, n64, t64 = relax_timed(s64, E0, tol, args.max_steps, args.check_every)
    print(f"[bench] fp64 form: {n64} steps, {t64:.2f} s, E={E64:.12f}, max|EG|={ea64:.2e}", file=sys.stderr,
          flush=True)
    F64 = s64.get_field()
    Q64 = s64.observables()["skyrmion_number"]
    s64.close()

    form = args.form
    s32 = ssa.Simulation(p, th, precision=form)
    s32.initialize({"mode": "array", "field": F0})
    s32.autotune_tiles(reps=3)
    tiles32 = s32.tile_variants()
    # FP64 lead (--start fp64): run the FP64 form while it converges faster
    # than the FP32 form would, and hand over once a chunk shrinks max|EG| by
    # less than --lead-rate (the FP32 form's late-phase rate, ~2.7x per 10 000
    # of its steps at ~0.7x the FP64 step cost, is ~1.33x per 2000 FP64 steps'
    # time).  The arrest kills per chunk are reported alongside.
    nA, tA, killsA, lead_converged = 0, 0.0, [], False
    rates = []
    eap = np.inf
    if args.start == "fp64":
        s32.set_precision("fp64")
        s32.use_graphs(args.graph_steps)
        E = s32.compute_energy()
        while nA < args.max_steps:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            Eprev = E
            E, _, eap = s32.steps(E, args.lead_chunk, history=True)
            s32.synchronize()
            tA += time.perf_counter() - t0
            nA += args.lead_chunk
            kills = int(np.sum(np.diff(np.concatenate([[Eprev], s32.last_history[0]])) > 0))
            killsA.append(kills)
            if eap < tol:
                lead_converged = True
                break
            if nA > args.lead_chunk:
                rates.append(ea_prev / eap)
                if rates[-1] < args.lead_rate:
                    break
            ea_prev = eap
        print(f"[bench] fp64 lead: {nA} steps, {tA:.2f} s, max|EG|={eap:.2e}, kills per {args.lead_chunk} steps "
              f"{killsA}, max|EG| decay per chunk {[round(r, 2) for r in rates]}", file=sys.stderr, flush=True)
        if not lead_converged:
            s32.set_precision(form)
            s32.set_tile_variants(tiles32["variants"])
            s32.set_fused_stages(tiles32["fused23"], tiles32["fused41"])
    E = s32.compute_energy()
    Ep = E
    n32 = npol_total = nchecks = npol = 0
    t32 = tpol = 0.0
    chunk, checks = args.chunk, []
    fp64_tail = 0
    while not lead_converged:
        s32.use_graphs(args.graph_steps)
        E, _, ea32, dn, dt_ = _chunk(s32, E, min(chunk, args.max_steps - n32))
        n32 += dn
        t32 += dt_
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s32.set_precision("fp64")
        s32.use_graphs(args.graph_steps)
        Ep, eap, npol, _ = relax_timed(s32, s32.compute_energy(), tol, args.polish, 20)
        tpol += time.perf_counter() - t0
        npol_total += npol
        nchecks += 1
        print(f"[bench] {form} {n32} steps ({t32:.2f} s, max|EG| {ea32:.2e}) + polish {npol}: E={Ep:.12f}, "
              f"max|EG|={eap:.2e}", file=sys.stderr, flush=True)
        if eap < tol or n32 >= args.max_steps:
            break
        checks.append((n32 + npol_total, eap))
        if len(checks) >= 2 and checks[-1][1] > 0.7 * checks[-2][1]:
            # FP32 has stalled (its steps no longer resolve the remaining
            # increments): finish in FP64
            t0 = time.perf_counter()
            Ep, eap, ntail, _ = relax_timed(s32, Ep, tol, max(0, args.max_steps - n32), args.check_every)
            tpol += time.perf_counter() - t0
            npol_total += ntail
            fp64_tail = ntail
            print(f"[bench] fp64 tail: {ntail} steps: E={Ep:.12f}, max|EG|={eap:.2e}", file=sys.stderr, flush=True)
            break
        # next chunk: the steps the last two checks' geometric decay predicts, within [chunk/10, chunk]
        if len(checks) >= 2 and checks[-1][1] < checks[-2][1]:
            (n0, e0), (n1, e1) = checks[-2], checks[-1]
            need = np.log(tol / e1) / (np.log(e1 / e0) / (n1 - n0))
            chunk = int(min(max(1.05 * need, args.chunk / 10), args.chunk))
        else:
            chunk = args.chunk
        t0 = time.perf_counter()
        s32.set_precision(form)
        s32.set_tile_variants(tiles32["variants"])
        tpol += time.perf_counter() - t0
        E = Ep
    Fp = s32.get_field()
    Qp = s32.observables()["skyrmion_number"]
    s32.close()
    dE = abs(Ep - E64) / abs(E64)
    dm = float(np.abs(Fp[:3] - F64[:3]).max())
    t_mixed = tA + t32 + tpol
    out = {
        "metric": f"time-to-converge, {form} form + FP64 polish vs FP64 form",
        "value": t64 / t_mixed,
        "unit": "speedup (FP64 form time / FP32+polish time)",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"chiral magnet {n}x{n}, h=10/319, Heusler DMI, {'demag' if args.demag else 'no demag'}, "
                               f"{args.nsk} seeded anti-skyrmion(s) R~U(0.3, 0.6), noise N(0, 1e-3) (seed "
                               f"{args.seed}); relax to max|EG| < {tol:g}",
                   "lattice": [n, n], "fp32_chunk_steps": args.chunk, "max_polish_steps": args.polish},
        "converged": bool(ea64 < tol and eap < tol),
        "fp64_form": {"steps": n64, "seconds": t64, "energy": E64, "err_absmax": ea64},
        "fp64_lead": {"steps": nA, "seconds": tA, "kills_per_chunk": killsA, "decay_per_chunk": rates,
                      "chunk": args.lead_chunk, "hand_over_rate": args.lead_rate,
                      "converged": lead_converged} if args.start == "fp64" else None,
        "fp32_phase": {"form": form, "steps": n32, "seconds": t32, "polish_checks": nchecks, "tiles": tiles32},
        "fp64_polish": {"steps_total": npol_total, "steps_final": npol, "tail_steps": fp64_tail, "seconds": tpol,
                        "energy": Ep, "err_absmax": eap},
        "skyrmion_number": {"fp64_form": Q64, "fp32_polished": Qp, "identical_rounded": round(Q64) == round(Qp)},
        "relE_vs_fp64_form": dE,
        "max_dm_vs_fp64_form": dm,
        "target_relE": 1e-7,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def converge_workload(args):
    """Time-to-converge of the two small FP64 workloads on the GPU: the
    Abelian-Higgs vortex, 128^2, 2000 steps (energy -> pi), and the chiral
    magnet, 320^2, relaxed to max |EG| < 1e-8 (the demag saddle makes this
    ~2e5 steps).  Autotuned shapes, graphs of 20 steps."""
    import soliton_solver_amd as ssa
    import torch
    cases = [
        ("abelian_higgs_128", "Abelian Higgs",
         dict(xlen=128, ylen=128, xsize=0.2 * 127, ysize=0.2 * 127, q=1.0, Lambda=1.0, u1=1.0, vortex_number=1.0,
              time_step=0.1),
         {"mode": "vortex", "noise": 1e-3, "seed": 0}, 0.0, 2000),
        ("chiral_magnet_320", "Chiral magnet",
         dict(xlen=320, ylen=320, xsize=10.0, ysize=10.0, J=40e-12, K=0.8e6, D=4e-3, M=580e3, B=0.0,
              dmi_term="Heusler", demag=True, unit_magnetization=True),
         {"mode": "skyrmions", "ansatz": "anti", "skyrmions": [(5.0, 5.0, 1.0)], "noise": 1e-6, "seed": 1},
         1e-8, 400000),
    ]
    res = {}
    for key, name, kw, cfg, tol, max_steps in cases:
        if args.converge_cap:  # a bounded run (profiling): at most this many steps per case
            max_steps = min(max_steps, args.converge_cap)
        th = ssa.load_theory(name)
        sim = ssa.Simulation(th.params.default_params(**kw), th, precision="fp64")
        sim.initialize(cfg)
        tiles = sim.autotune_tiles(reps=5)
        sim.use_graphs(20)
        E0 = sim.compute_energy()
        sim.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if tol > 0.0:
            E, ea, n = sim.minimize(tol=tol, max_steps=max_steps, check_every=5000, prev_energy=E0)
        else:
            E, _, ea = sim.steps(E0, max_steps)
            n = max_steps
        sim.synchronize()
        dt = time.perf_counter() - t0
        obs = sim.observables()
        charge = {k: v for k, v in obs.items() if k != "energy"}
        res[key] = {"seconds": dt, "steps": n, "us_per_step": dt / n * 1e6, "energy": E, "err_absmax": ea,
                    "tol": tol, "charges": charge, "tile_variants": tiles["variants"], "fused23": tiles["fused23"],
                    "fused41": tiles["fused41"]}
        print(f"[bench] {key}: {n} steps in {dt:.2f} s ({dt / n * 1e6:.1f} us/step), E={E:.12f}, "
              f"max|EG|={ea:.2e}, {charge}", file=sys.stderr, flush=True)
        sim.close()
    out = {
        "metric": "time-to-converge, chiral magnet 320^2 FP64 to max|EG| < 1e-8",
        "value": res["chiral_magnet_320"]["seconds"],
        "unit": "s",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "chiral magnet 320x320, xsize 10, Heusler DMI, demag, one anti-skyrmion of radius 1 "
                               "+ N(0, 1e-6); Abelian-Higgs vortex 128x128 spacing 0.2, 2000 steps beside it"},
        "runs": res,
    }
    print(json.dumps(out), flush=True)


def numpy_oracle_rates(gl_steps, cm_steps):
    """The NumPy oracle (single thread), first steps of the two small reference
    workloads: Abelian Higgs 128^2 (spacing 0.2, dt 0.1, tanh(r) e^{i theta}
    vortex + U(-1e-3, 1e-3)) and chiral magnet 320^2 (xsize 10, Heusler,
    demag, one anti-skyrmion of radius 1 + N(0, 1e-6))."""
    import soliton_solver_amd as ssa
    from oracle import core as ocore
    from oracle.theories import chiral_magnet as o_cm
    from oracle.theories import ginzburg_landau as o_gl
    out = {}
    cases = [
        ("abelian_higgs_128", "Abelian Higgs", o_gl, gl_steps,
         dict(xlen=128, ylen=128, xsize=0.2 * 127, ysize=0.2 * 127, q=1.0, Lambda=1.0, u1=1.0, vortex_number=1.0,
              time_step=0.1),
         {"mode": "vortex", "noise": 1e-3, "seed": 0}),
        ("chiral_magnet_320", "Chiral magnet", o_cm, cm_steps,
         dict(xlen=320, ylen=320, xsize=10.0, ysize=10.0, J=40e-12, K=0.8e6, D=4e-3, M=580e3, B=0.0,
              dmi_term="Heusler", demag=True, unit_magnetization=True),
         {"mode": "skyrmions", "ansatz": "anti", "skyrmions": [(5.0, 5.0, 1.0)], "noise": 1e-6, "seed": 1}),
    ]
    for key, name, o, steps, kw, cfg in cases:
        if steps <= 0:
            continue
        th = ssa.load_theory(name)
        p_i, p_f = th.params.pack_device_params(th.params.default_params(**kw).resolved())
        p_i = np.ascontiguousarray(p_i, dtype=np.int32)
        p_f = np.ascontiguousarray(p_f, dtype=np.float64)
        X, Y, nf = int(p_i[0]), int(p_i[1]), int(p_i[4])
        g = np.empty((2, X, Y))
        g[0] = p_f[2] * np.arange(X)[:, None]
        g[1] = p_f[3] * np.arange(Y)[None, :]
        F = np.zeros((nf, X, Y))
        V = np.zeros_like(F)
        th.initial_config.initialize(Field=F, Velocity=V, grid=g, p_i_h=p_i, p_f_h=p_f, config=cfg)
        E0 = ocore.theory_energy(o, F, p_i, p_f)
        t0 = time.perf_counter()
        ocore.relax(o, F, V, p_i, p_f, steps, prev_energy=E0)
        dt = time.perf_counter() - t0
        out[key] = {"steps": steps, "seconds": dt, "updates_per_s": X * Y * steps / dt}
    return out


# rocprofv3 byte counters against a plain 1 GiB device copy on MI355X
# (bench.py --copy-calibration): FETCH_SIZE reports 0.5 KiB per KiB read,
# WRITE_SIZE 1.0 KiB per KiB written.
FETCH_SCALE = 2.0
WRITE_SCALE = 1.0


TRAFFIC_STEPS = 6


def measure_traffic(child_args, kernel):
    """HBM bytes per launch of ``kernel`` (a kernel name as rocprof prints it,
    without namespace and arguments): two ``rocprofv3 --pmc`` runs (FETCH_SIZE,
    then WRITE_SIZE) of a child ``bench.py --traffic-child`` that runs
    TRAFFIC_STEPS eager steps with the same form and shapes, each scaled by the
    copy calibration, averaged over the child's launches of the kernel (one per
    step).  Also returns which of the child's steps ran arrested, so that its
    bytes are compared with the algorithmic bytes of the same launches.  None
    when rocprofv3 is absent or a run fails."""
    import glob
    import re
    import shutil
    import sqlite3
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    kib = {}
    arrested = None
    with tempfile.TemporaryDirectory() as d:
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(d, ctr.lower())
            cmd = ["timeout", "-s", "KILL", "150", prof, "--pmc", ctr, "-d", out, "-o", "t", "--", sys.executable,
                   os.path.abspath(__file__), "--traffic-child", str(TRAFFIC_STEPS), *child_args]
            print(f"[bench] traffic: rocprofv3 --pmc {ctr} ({TRAFFIC_STEPS} steps)", file=sys.stderr, flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True, cwd=ROOT)
            dbs = glob.glob(os.path.join(out, "**", "*.db"), recursive=True)
            marks = [ln for ln in r.stdout.splitlines() if ln.startswith("TRAFFIC_CHILD ")]
            if r.returncode != 0 or not dbs or not marks:
                print(f"[bench] traffic: rocprofv3 run failed (rc {r.returncode})", file=sys.stderr, flush=True)
                return None
            pattern = json.loads(marks[-1].split(" ", 1)[1])["arrested"]
            if arrested is not None and pattern != arrested:
                print("[bench] traffic: the two counter runs took different arrest decisions", file=sys.stderr,
                      flush=True)
                return None
            arrested = pattern
            con = sqlite3.connect(dbs[0])
            rows = con.execute("select kernel_name, count(value), avg(value) from counters_collection "
                               "where counter_name = ? group by kernel_name", (ctr,)).fetchall()
            con.close()
            for name, cnt, v in rows:
                short = re.sub(r"\(.*$", "", re.sub(r"^void ", "", name)).replace("ss::", "")
                if short == kernel and cnt == len(arrested):
                    kib[ctr] = float(v)
                elif short == kernel:
                    print(f"[bench] traffic: {kernel} ran {cnt} times in the child, not once per step",
                          file=sys.stderr, flush=True)
    if len(kib) != 2:
        return None
    return {"bytes": kib["FETCH_SIZE"] * FETCH_SCALE * 1024 + kib["WRITE_SIZE"] * WRITE_SCALE * 1024,
            "read_bytes": kib["FETCH_SIZE"] * FETCH_SCALE * 1024, "write_bytes": kib["WRITE_SIZE"] * WRITE_SCALE * 1024,
            "arrested": arrested}


def schedule(tiles, comp):
    """Words per site and component of the launch that finishes a step (the
    dominant one, timed as the pass-4 slot of ``time_passes``) and of a whole
    step, for the schedule ``tiles`` describes (the dict of
    ``Simulation.tile_variants``); ``comp``: the compensated FP32 form, whose
    field carries a second word (lo) that its pass 1 reads and its combine
    reads and writes.  ``arrested_words``: the dominant launch's words in a
    step whose velocity was arrested, where V is known to be zero and is not
    read (and a paired 3+4 sweep takes l2 = l1, read once)."""
    pairs, f41, f23 = bool(tiles["pairs"]), bool(tiles["fused41"]), bool(tiles["fused23"])
    v4 = tiles["variants"][3]
    cw = 2 if v4 >= 6 else 1
    if comp:
        if pairs and tiles.get("comp12"):  # 1+2 (F, lo, V -> l1, l2) + 3+4 (F, lo, V, l1, l2 -> F', lo', V')
            return {"name": "compensated paired 1+2, 3+4", "dominant_words": 8, "arrested_words": 6,
                    "step_words": 13, "kernel": f"stream_pass_comp<ChiralMagnet, float, 34, {cw}>"}
        if pairs:  # pass 1 (F, lo -> l1) + pass 2 (F, V -> l2) + 3+4 (F, lo, V, l1, l2 -> F', lo', V')
            return {"name": "compensated pass 1, pass 2, paired 3+4", "dominant_words": 8, "arrested_words": 6,
                    "step_words": 14, "kernel": f"stream_pass_comp<ChiralMagnet, float, 34, {cw}>"}
        kern = (f"stream_pass_comp<ChiralMagnet, float, 4, {cw}>" if v4 >= 3 else
                f"stage_pass<ChiralMagnet, float, 4, {tiles['rows'][3]}, true>")
        return {"name": "compensated pass 1, " + ("2+3, " if f23 else "2, 3, ") + "4",
                "dominant_words": 9, "arrested_words": 8, "step_words": 3 + (5 if f23 else 7) + 9, "kernel": kern}
    if pairs:  # 1+2 (F, V -> l1, l2) + 3+4 (F, V, l1, l2 -> F', V')
        return {"name": "paired 1+2, 3+4", "dominant_words": 6, "arrested_words": 4, "step_words": 10,
                "kernel": f"stream_pass<ChiralMagnet, float, 34, {cw}>"}
    if f41:  # 2+3 or 2, 3, then 4+1 (F, V, l1, l2, l3 -> F', V', l1')
        return {"name": ("2+3, " if f23 else "2, 3, ") + "4+1", "dominant_words": 9, "arrested_words": 8,
                "step_words": (5 if f23 else 7) + 9, "kernel": f"stream_pass<ChiralMagnet, float, 41, {cw}>"}
    kern = (f"stream_pass<ChiralMagnet, float, 4, {cw}>" if v4 >= 3 else
            f"stage_pass<ChiralMagnet, float, 4, {tiles['rows'][3]}>")
    return {"name": "1, " + ("2+3, " if f23 else "2, 3, ") + "4", "dominant_words": 7, "arrested_words": 6,
            "step_words": 2 + (5 if f23 else 7) + 7, "kernel": kern}


def launch_words(sched, kills):
    """Mean algorithmic words per site and component of the dominant launch
    over steps whose arrest decisions are ``kills``: the schedule's
    ``dominant_words`` in a step with the velocity free, ``arrested_words`` in
    an arrested one (V known to be zero and not read)."""
    k = np.asarray(kills, dtype=bool)
    return float(((~k).sum() * sched["dominant_words"] + k.sum() * sched["arrested_words"]) / len(k))


def arrested_steps(energies, prev_energy):
    """Which steps of a call ran with the velocity arrested: the arrest of step
    s >= 1 is E(F_s) > E(F_{s-1}) (core/integrator.py:338-343), with E(F_0) the
    call's prev_energy; ``energies[k]`` = E(F_{k+1}), the per-step history of
    the call.  Step 0 is never arrested (the decision the call starts from is
    applied before it, on the host's side)."""
    e = np.asarray(energies, dtype=np.float64)
    prevs = np.concatenate([[prev_energy], e[:-1]])
    kill = np.zeros(len(e), dtype=bool)
    kill[1:] = e[:-1] > prevs[:-1]
    return kill


def traffic_child(th, p, cfg, form, tiles, args):
    """The counted child of ``measure_traffic``: ``args.traffic_child`` eager
    steps of ``form`` from the seeded field, with the launch shapes ``tiles``
    and the schedule of --fused23 / --fused41 / --pairs; prints which of them
    ran arrested."""
    import soliton_solver_amd as ssa
    sim = ssa.Simulation(p, th, precision=form)
    sim.initialize(cfg)
    if tiles is not None:
        sim.set_tile_variants(tiles)
        sim.set_fused_stages(args.fused23 > 0, args.fused41 > 0, pairs=args.pairs > 0)
    E0 = sim.compute_energy()
    sim.steps(E0, args.traffic_child)
    energies = sim.history(args.traffic_child)[0]
    kills = arrested_steps(energies, E0)
    sim.close()
    print("TRAFFIC_CHILD " + json.dumps({"arrested": [bool(k) for k in kills], "E0": E0,
                                         "energies": [float(e) for e in energies]}), flush=True)


def cpu_baseline(args, n):
    """The oracle's C++/OpenMP FP64 restatement of the reference path (unfused,
    derivative scratch) on a bounded sample of the chiral workload, plus the
    single-thread NumPy oracle on the first steps of the two small workloads."""
    import soliton_solver_amd as ssa
    from oracle import cpu_build
    m = args.cpu_sample
    thc, pc, cfgc = chiral_workload(m, seed=3, nsk=max(1, round(8 * (m / n) ** 2)))
    simc = ssa.Simulation(pc, thc)
    simc.initialize(cfgc)
    Fc = simc.get_field().copy()
    Vc = np.zeros_like(Fc)
    p_i, p_f = simc.p_i_h, simc.p_f_h
    simc.close()
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    t0 = time.perf_counter()
    cpu_build.chiral_relax(p_i, p_f, Fc, Vc, np.inf, args.cpu_steps, nthreads=threads)
    tc = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cpu = {"value": m * m * args.cpu_steps / tc, "unit": "grid-point updates/s", "cores": threads,
           "kind": "port",
           "sample": f"C++/OpenMP FP64 restatement of the reference path (unfused, derivative scratch), "
                     f"chiral magnet {m}x{m} same parameters, {args.cpu_steps} steps, {tc:.1f} s; "
                     f"host {model or platform.processor()}, os.cpu_count()={os.cpu_count()}, "
                     f"{threads} threads (the process's CPU affinity)"}
    gl_steps, cm_steps = (int(v) for v in args.numpy_steps.split(","))
    cpu["numpy_oracle_single_thread"] = numpy_oracle_rates(gl_steps, cm_steps)
    print(f"[bench] numpy oracle: {cpu['numpy_oracle_single_thread']}", file=sys.stderr, flush=True)
    return cpu


def _chunk(sim, E, nsteps):
    """``nsteps`` steps in one timed call; returns (E, err, err_absmax, steps, seconds)."""
    import torch
    sim.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E, err, ea = sim.steps(E, nsteps)
    sim.synchronize()
    return E, err, ea, nsteps, time.perf_counter() - t0


def make_batch(seed, ninst):
    """The batched workload's instances (FP32, initialised, launch shape
    autotuned): sizes uniform in {64, 128, 256, 512}^2, xsize 20, dt = 0.0957 h,
    pion mass uniform in [0.1, 1], Skyrme coupling uniform in [0.5, 2],
    standard potential, Neel hedgehog of charge 1..4 plus U(-1e-3, 1e-3).
    Returns (batch, sizes, charges, autotune info, pion masses, Skyrme couplings)."""
    import soliton_solver_amd as ssa
    from soliton_solver_amd.batch import BatchSimulation
    th = ssa.load_theory("Baby Skyrme model")
    rng = np.random.default_rng(seed)
    sizes = rng.choice([64, 128, 256, 512], size=ninst)
    mpi = rng.uniform(0.1, 1.0, size=ninst)
    kappa = rng.uniform(0.5, 2.0, size=ninst)
    charge = rng.integers(1, 5, size=ninst)
    params = []
    for n, m, k in zip(sizes, mpi, kappa):
        h = 20.0 / (int(n) - 1)
        params.append(th.params.default_params(xlen=int(n), ylen=int(n), xsize=20.0, ysize=20.0, mpi=float(m),
                                               kappa=float(k), unit_magnetization=True, time_step=0.0957 * h))
    bs = BatchSimulation(th, params, precision="fp32")
    bs.initialize([{"mode": "initial", "ansatz": "neel", "skyrmion_number": int(q), "noise": 1e-3, "seed": i}
                   for i, q in enumerate(charge)])
    tinfo = bs.autotune_tiles(reps=5)
    return bs, sizes, charge, tinfo, mpi, kappa


def batch_throughput(args, steps=100, warmup=20):
    """The batched workload's step rate with every instance active (tolerance
    0, so none retires): ``warmup`` steps (the step graphs are captured there),
    then ``steps`` timed steps in graphs of 20."""
    import torch
    bs, sizes, charge, tinfo, _, _ = make_batch(args.seed, args.instances)
    bs.use_graphs(20)
    bs.begin()
    bs.run(warmup, 0.0, check_every=warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bs.run(steps, 0.0, check_every=steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    sites = float(np.sum(np.asarray(sizes, dtype=np.float64) ** 2))
    out = {"workload": f"{args.instances} baby Skyrme lattices, sizes U{{64,128,256,512}}^2 (seed {args.seed}), FP32 "
                       "fields with FP64 sums, every instance active",
           "updates_per_s": sites * steps / dt, "ms_per_step": dt / steps * 1e3, "steps": steps,
           "total_sites": int(sites), "shape_variant": tinfo["variant"], "fused_passes_2_3": bool(tinfo["fused23"]),
           "paired_sweeps": bool(tinfo.get("pairs", False)),
           "relaxation": "bench.py --workload batch (until max|EG| < 1e-6 or a step budget)"}
    bs.close()
    return out


def batch_workload(args):
    """Many baby-Skyrme relaxations of uneven sizes in one set of launches per RK4
    pass: sizes uniform in {64, 128, 256, 512}^2, xsize 20, dt = 0.0957 h, pion
    mass uniform in [0.1, 1], Skyrme coupling uniform in [0.5, 2], standard
    potential, Neel hedgehog of charge 1..4 plus U(-1e-3, 1e-3), graphs of 20
    steps.  FP32 fields with FP64 sums until every instance reaches max |EG| <
    1e-3 (the FP32 floor) or --fp32-steps, then, for a tighter --tol, the FP64
    retire phase (fields converted in place, every instance reactivated) until
    each reaches --tol, within --batch-max-steps steps in all."""
    import torch
    ninst = args.instances
    bs, sizes, charge, tinfo, mpi, kappa = make_batch(args.seed, ninst)
    print(f"[bench] batch tile variant {tinfo['variant']} ({tinfo['rows']} rows), fused 2+3: {tinfo['fused23']}; "
          "ms per variant " + " ".join(f"{x:.3f}" for x in tinfo["ms"]) + "; with 2+3 fused "
          + " ".join(f"{x:.3f}" for x in tinfo["ms_fused23"]) + "; paired sweeps "
          + " ".join(f"{x:.3f}" for x in tinfo["ms_pairs"]) + f" (chosen: {tinfo['pairs']})",
          file=sys.stderr, flush=True)
    v = tinfo["variant"]
    pass_ms = float(tinfo["ms_pairs"][v] if tinfo["pairs"] else
                    tinfo["ms_fused23"][v] if tinfo["fused23"] else tinfo["ms"][v])
    bs.use_graphs(20)
    sites = np.array([int(n) * int(n) for n in sizes], dtype=np.float64)
    phase_t = {}

    def report(prec, done, active, st):
        phase_t.setdefault(prec, time.perf_counter())
        print(f"[bench]   {prec} {done} steps: {active} of {ninst} instances active, "
              f"max err {float(np.nanmax(st['err_absmax'])):.2e}", file=sys.stderr, flush=True)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = bs.relax(args.tol, args.batch_max_steps, check_every=args.check_every, reduced_tol=1e-3,
                   reduced_steps=args.fp32_steps, on_chunk=report)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    updates = float((sites * (res["steps_reduced"] + res["steps_fp64"])).sum())
    conv = res["converged_step"]
    by_charge = {int(q): {"instances": int((charge == q).sum()), "converged": int(((charge == q) & (conv >= 0)).sum()),
                          "median_err_absmax": float(np.median(res["err_absmax"][charge == q]))}
                 for q in range(1, 5)}
    by_size = {int(n): {"instances": int((sizes == n).sum()), "converged": int(((sizes == n) & (conv >= 0)).sum())}
               for n in (64, 128, 256, 512)}
    mbins = ((0.1, 0.4), (0.4, 0.7), (0.7, 1.0))
    by_mass = {f"{lo:.1f}-{hi:.1f}": {q: {"instances": int(((mpi >= lo) & (mpi < hi) & (charge == q)).sum()),
                                           "converged": int(((mpi >= lo) & (mpi < hi) & (charge == q)
                                                             & (conv >= 0)).sum())}
                                      for q in range(1, 5)}
               for lo, hi in mbins}
    out = {
        "metric": "grid-pt updates/s, batched relaxations (sum over instances of sites x steps / wall time)",
        "value": updates / wall,
        "unit": "grid-point updates/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32+f64" if args.tol < 1e-3 else "f32",
        "data": "synthetic",
        "config": {"workload": f"{ninst} baby Skyrme relaxations, sizes U{{64,128,256,512}}^2, seed {args.seed}, "
                               f"FP32 to max|EG| < 1e-3 (<= {args.fp32_steps} steps), then FP64 to max|EG| < "
                               f"{args.tol:g}, <= {args.batch_max_steps} steps in all",
                   "instances": ninst, "total_sites": int(sites.sum()), "tiles": bs.num_tiles,
                   "tile_rows": tinfo["rows"], "shape_variant": v, "fused_passes_2_3": bool(tinfo["fused23"]),
                   "paired_sweeps": bool(tinfo.get("pairs", False))},
        "passes_ms_all_active_fp32": pass_ms,
        "all_active_updates_per_s_fp32": float(sites.sum()) / (pass_ms * 1e-3),
        "time_to_converge_s": wall,
        "converged": res["converged"],
        "failed": res["failed"],
        "converged_by_charge": by_charge,
        "converged_by_size": by_size,
        "converged_by_pion_mass_and_charge": by_mass,
        "per_instance_note": "size, charge, pion mass, Skyrme coupling, step converged at (-1: not within the "
                             "budget, -2: failed), last max|EG|",
        "per_instance": [[int(n), int(q), round(float(m), 4), round(float(k), 4), int(c), float(f"{e:.3e}")]
                         for n, q, m, k, c, e in zip(sizes, charge, mpi, kappa, conv, res["err_absmax"])],
        "steps_fp32_max": int(res["steps_reduced"].max()),
        "steps_fp64_max": int(res["steps_fp64"].max()),
    }
    bs.close()
    print(json.dumps(out), flush=True)


def anisotropic_betas(seed):
    """beta1..beta4 ~ U[0.2, 1] from ``seed`` (alpha1 = -1, alpha2 = -0.5),
    redrawn until the quartic is bounded below with real, non-zero vacua."""
    rng = np.random.default_rng(seed)
    while True:
        b1, b2, b3, b4 = rng.uniform(0.2, 1.0, size=4)
        b3t = b3 - 2.0 * b4
        den = b3t * b3t - 4.0 * b1 * b2
        n1, n2 = 2.0 * b2 * -1.0 - b3t * -0.5, 2.0 * b1 * -0.5 - b3t * -1.0
        if b3t > -2.0 * np.sqrt(b1 * b2) + 0.1 and den != 0 and n1 / den > 0 and n2 / den > 0:
            return float(b1), float(b2), float(b3), float(b4)


def differentiation_workload(args):
    """Anisotropic s+id Ginzburg-Landau, 512 x 512, FP64, the exact-adjoint
    flow (the variation is dE_h/dphi of the discrete energy): alpha1 = -1,
    alpha2 = -0.5, beta1..beta4 ~ U[0.2, 1] from seed 1 (redrawn until the
    quartic is bounded below), gamma1 = gamma2 = 1, gamma12 = 0.3, kappa =
    0.673274, Q = 1; vortex ansatz (N1 = N2 = 1) plus U(-1e-2, 1e-2).  Relax to
    max |dE_h/dphi|/gv < --tol, then dE*/dbeta_i by the backward of the
    discrete energy (exact: dual numbers in the packed couplings, chained
    through ResolvedParams) against central re-relaxations at beta_i +- 1e-3,
    and torch.autograd.gradcheck of the energy on a 32 x 32 patch."""
    import torch
    import soliton_solver_amd as ssa
    from soliton_solver_amd import autograd as sag
    th = ssa.load_theory("Anisotropic superconductor")
    n, xs = 512, args.diff_xsize
    h = xs / (n - 1)
    b1, b2, b3, b4 = anisotropic_betas(1)
    p = th.params.default_params(xlen=n, ylen=n, xsize=xs, ysize=xs, alpha1=-1.0, alpha2=-0.5, beta1=b1, beta2=b2,
                                 beta3=b3, beta4=b4, gamma1=1.0, gamma2=1.0, gamma12=0.3, Q=1.0, kappa=0.673274,
                                 time_step=args.diff_dt * h)
    sim = ssa.Simulation(p, th, variation="exact_adjoint")
    sim.initialize({"mode": "initial", "noise": 1e-2, "seed": 1})
    p_f = sim.p_f_h
    sim.use_graphs(20)
    t0 = time.perf_counter()
    # legs of 20 000 steps; a leg that shrinks max |EG| by less than 2x restarts the next one from V = 0
    E, nrel, restarts, prev = sim.compute_energy(), 0, 0, float("inf")
    while nrel < args.diff_max_steps:
        E, ea, n = sim.minimize(tol=args.tol, max_steps=min(20000, args.diff_max_steps - nrel), check_every=5000,
                                prev_energy=E)
        nrel += n
        if ea < args.tol:
            break
        if ea > 0.5 * prev:
            sim.set_velocity(np.zeros(sim.field_shape))
            E = sim.compute_energy()
            restarts += 1
        prev = ea
    t_relax = time.perf_counter() - t0
    Fstar = sim.get_field()
    field = torch.tensor(Fstar, device="cuda")
    G = sag.exact_gradient(sim, field)
    gmax = float(G.abs().max()) / float(p_f[4])
    print(f"[bench] 512^2 exact flow (xsize {xs}, dt {args.diff_dt} h): {nrel} steps, {restarts} restarts from rest, "
          f"{t_relax:.1f} s, E*={E:.13f}, "
          f"max|EG|={ea:.2e}, max|dE/dphi|/gv={gmax:.2e}", file=sys.stderr, flush=True)
    names = ["beta1", "beta2", "beta3", "beta4"]
    t0 = time.perf_counter()
    grads = sag.energy_parameter_gradient(sim, field, names)
    t_back = time.perf_counter() - t0
    sim.close()
    sens = {}
    for name in names:
        v = float(getattr(p, name))
        d = 1e-3 * max(abs(v), 1.0)
        Es = {}
        steps = []
        for dd in (d, d / 2):  # central re-relaxations at +-d and +-d/2
            for sgn in (+1, -1):
                s2 = ssa.Simulation(p.with_(**{name: v + sgn * dd}), th, variation="exact_adjoint")
                s2.initialize({"mode": "array", "field": Fstar})
                s2.use_graphs(20)
                E2, ea2, n2 = s2.minimize(tol=args.tol, max_steps=args.diff_max_steps, check_every=5000)
                Es[(dd, sgn)] = E2
                steps.append([n2, ea2])
                s2.close()
        fd = (Es[(d, 1)] - Es[(d, -1)]) / (2 * d)
        fd_half = (Es[(d / 2, 1)] - Es[(d / 2, -1)]) / d
        # Richardson: the O(d^2) truncation of the central difference cancels
        fd_rich = (4.0 * fd_half - fd) / 3.0
        rel = abs(grads[name] - fd) / abs(fd)
        rel_rich = abs(grads[name] - fd_rich) / abs(fd_rich)
        sens[name] = {"backward": grads[name], "rerelaxation": fd, "rel_diff": rel,
                      "rerelaxation_richardson": fd_rich, "rel_diff_richardson": rel_rich,
                      "rerelax_steps_err": steps}
        print(f"[bench] dE*/d{name}: backward {grads[name]:.10e} re-relaxation {fd:.10e} rel {rel:.2e}; "
              f"Richardson (d, d/2) {fd_rich:.10e} rel {rel_rich:.2e}", file=sys.stderr, flush=True)
    # gradcheck of the energy's backward against central differences on a 32 x 32 patch
    sim = ssa.Simulation(p, th)
    base = torch.tensor(Fstar, device="cuda")
    x0, y0 = 240, 250
    patch = base[:, x0:x0 + 32, y0:y0 + 32].clone().requires_grad_(True)

    def f(pt):
        full = base.clone()
        full[:, x0:x0 + 32, y0:y0 + 32] = pt
        return sag.discrete_energy(full, sim)

    ok = bool(torch.autograd.gradcheck(f, (patch,), eps=1e-6, atol=1e-6, rtol=1e-4, nondet_tol=0.0,
                                       fast_mode=False))
    sim.close()
    out = {
        "metric": "dE*/dbeta_i at a relaxed minimum: backward of the discrete energy vs re-relaxation "
                  "(central differences at d and d/2, Richardson-combined)",
        "value": max(v["rel_diff_richardson"] for v in sens.values()),
        "unit": "max relative difference",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"anisotropic s+id 512x512, xsize {xs}, exact-adjoint flow dt {args.diff_dt} h, "
                               f"betas from seed 1, vortex ansatz + U(-1e-2, 1e-2), relaxed to {args.tol:g}"},
        "relaxation": {"steps": nrel, "restarts_from_rest": restarts, "seconds": t_relax, "energy": E, "err_absmax": ea,
                       "max_dEdphi_over_gv": gmax, "converged": bool(ea < args.tol)},
        "backward_seconds": t_back,
        "sensitivities": sens,
        "gradcheck_32x32_patch": ok,
    }
    print(json.dumps(out), flush=True)


# The nine theories as the theories workload runs them: registry name, oracle
# module (the CPU baseline), parameters besides the lattice, box side, time
# step in units of the spacing (None: the theory's default 0.5) and initial
# condition.
THEORY_CASES = (
    ("Chiral magnet", "chiral_magnet",
     dict(J=40e-12, K=0.8e6, D=4e-3, M=580e3, B=0.0, dmi_term="Heusler", demag=True, unit_magnetization=True),
     10.0, None, {"mode": "initial", "ansatz": "anti", "noise": 1e-3}),
    ("Ginzburg-Landau superconductor", "ginzburg_landau", dict(q=1.0, Lambda=1.0, u1=1.0, vortex_number=1.0),
     25.4, None, {"mode": "vortex", "noise": 1e-3}),
    ("Baby Skyrme model", "baby_skyrme", dict(mpi=1.0, kappa=1.0, unit_magnetization=True),
     20.0, 0.0957, {"mode": "initial", "ansatz": "neel", "skyrmion_number": 2, "noise": 1e-3}),
    ("Anisotropic superconductor", "anisotropic_superconductor",
     dict(alpha1=-1.0, alpha2=-0.5, gamma1=1.0, gamma2=1.0, gamma12=0.3, Q=1.0, kappa=0.673274),
     25.0, 0.3, {"mode": "initial", "noise": 1e-2}),
    ("Bose-Einstein condensate", "bose_einstein_condensate", dict(beta=50.0, omega_rot=0.5, vortex_number=2.0),
     12.0, None, {"mode": "ground", "noise": 1e-3}),
    ("Anyon superconductor", "maxwell_chern_simons_higgs", dict(q=1.0, Lambda=1.5, kappa=0.7, u1=1.0),
     12.0, None, {"mode": "initial", "noise": 1e-2}),
    ("Ferromagnetic superconductor", "ferromagnetic_superconductor",
     dict(q=0.8, eta1=-0.1, gamma=1.2, unit_magnetization=True),
     16.0, None, {"mode": "initial", "ansatz": "bloch", "vortex_number": 1.0, "noise": 1e-2}),
    ("Spin triplet superconducting ferromagnet", "spin_triplet", dict(q=0.9, gamma=1.1, unit_magnetization=True),
     16.0, None, {"mode": "initial", "ansatz": "bloch", "noise": 1e-2}),
    ("Liquid crystal", "liquid_crystal", dict(unit_magnetization=True),
     12.0, None, {"mode": "initial", "ansatz": "bloch", "noise": 1e-2}),
)


def theory_params(th, name, kw, xsize, dt_h, n):
    extra = dict(kw)
    if name == "Anisotropic superconductor":
        extra.update(zip(("beta1", "beta2", "beta3", "beta4"), anisotropic_betas(1)))
    if dt_h is not None:
        extra["time_step"] = dt_h * xsize / (n - 1)
    return th.params.default_params(xlen=n, ylen=n, xsize=xsize, ysize=xsize, **extra)


def theories_workload(args):
    """Every theory on one n x n lattice (--theories-n, default 2048): the FP64
    form and the reduced form the theory has (compensated FP32, else plain
    FP32; the L2-normalised BEC flow has FP64 storage only), each autotuned,
    --warmup steps then --steps timed in graphs of 20 -> updates/s and the
    fraction of 8 TB/s at the fixed 16 nf w bytes per update; the launch that
    finishes a step timed by HIP events over the 20 steps that follow and
    priced at its algorithmic words with those steps' arrest decisions; and,
    beside it, the single-thread NumPy oracle on a 128 x 128 lattice of the
    same theory (--theories-cpu-steps steps)."""
    import importlib
    import torch
    import soliton_solver_amd as ssa
    from oracle import core as ocore
    n, steps, warmup = args.theories_n, args.steps, args.warmup
    N = n * n
    out = {}
    for name, oname, kw, xsize, dt_h, cfg in THEORY_CASES:
        th = ssa.load_theory(name)
        p = theory_params(th, name, kw, xsize, dt_h, n)
        cfg = dict(cfg, seed=1)
        forms = {}
        for form in ("fp64", "fp32c", "fp32"):
            if form == "fp32" and "refused" not in forms.get("fp32c", {"refused": ""}):
                continue  # the compensated form ran: the plain one is not needed
            try:  # a form the theory does not have is refused at creation or at its first step
                sim = ssa.Simulation(p, th, precision=form)
            except (ValueError, RuntimeError) as e:
                forms[form] = {"refused": str(e)}
                continue
            try:
                sim.initialize(cfg)
                try:
                    tiles = sim.autotune_tiles(reps=5)
                except ValueError:  # the L2-normalised flow has fixed launch shapes
                    tiles = sim.tile_variants()
                sim.use_graphs(args.graph_steps)
                E = sim.compute_energy()
                E, _, _ = sim.steps(E, warmup)
                sim.prepare_graphs(steps)
            except (ValueError, RuntimeError) as e:
                sim.close()
                forms[form] = {"refused": str(e)}
                continue
            nf = int(sim.p_i_h[4])
            w = 8 if form == "fp64" else 4
            sim.synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            E1, _, ea = sim.steps(E, steps)
            sim.synchronize()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            timed_kills = arrested_steps(sim.history(steps)[0], E)
            try:
                passes = sim.time_passes(E1, 20)
                kills = arrested_steps(sim.history(20)[0], E1)
            except ValueError:  # the L2-normalised flow is timed as a whole only
                passes, kills = np.zeros(6), None
            sim.close()
            ups = N * steps / dt
            res = {"ms_per_step": dt / steps * 1e3, "updates_per_s": ups, "nf": nf,
                   "hbm_frac_alg": ups * 16 * nf * w / (HBM_PEAK_GBS * 1e9),
                   "timed_steps_arrested": int(timed_kills.sum()), "energy_after": E1, "err_absmax_after": ea,
                   "tile_variants": tiles["variants"], "paired_sweeps": bool(tiles["pairs"]),
                   "pass_ms": [round(float(x), 5) for x in passes]}
            if kills is not None and passes[4] > 0:
                sched = schedule(tiles, form == "fp32c")
                words = launch_words(sched, kills)
                gbs = words * nf * w * N / (passes[4] * 1e-3) / 1e9
                res.update({"schedule": sched["name"], "words_per_update": sched["step_words"],
                            "hbm_frac_schedule": ups * sched["step_words"] * nf * w / (HBM_PEAK_GBS * 1e9),
                            "dominant_launch": {"ms": float(passes[4]), "alg_words_per_site": words * nf,
                                                "steps_arrested": f"{int(kills.sum())} of 20",
                                                "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS}})
            else:
                res["schedule"] = "L2-normalised flow: stage passes plus a normalisation of every stage field"
            forms[form] = res
            print(f"[bench] {name} {form}: {res['ms_per_step']:.3f} ms/step, {ups:.3e} updates/s, "
                  f"{100 * res['hbm_frac_alg']:.1f} % of 8 TB/s at 16 nf w"
                  + (f", finishing launch {res['dominant_launch']['achieved_gbs'] / 1e3:.2f} TB/s"
                     if "dominant_launch" in res else ""), file=sys.stderr, flush=True)
        # the single-thread NumPy oracle on a 128 x 128 lattice of the same theory
        o = importlib.import_module(f"oracle.theories.{oname}")
        pc = theory_params(th, name, kw, xsize, dt_h, 128)
        p_i, p_f = th.params.pack_device_params(pc.resolved())
        p_i = np.ascontiguousarray(p_i, dtype=np.int32)
        p_f = np.ascontiguousarray(p_f, dtype=np.float64)
        X, Y, nfc = int(p_i[0]), int(p_i[1]), int(p_i[4])
        g = np.empty((2, X, Y))
        g[0] = p_f[2] * np.arange(X)[:, None]
        g[1] = p_f[3] * np.arange(Y)[None, :]
        F = np.zeros((nfc, X, Y))
        V = np.zeros_like(F)
        th.initial_config.initialize(Field=F, Velocity=V, grid=g, p_i_h=p_i, p_f_h=p_f, config=cfg)
        E0 = ocore.theory_energy(o, F, p_i, p_f)
        t0 = time.perf_counter()
        ocore.relax(o, F, V, p_i, p_f, args.theories_cpu_steps, prev_energy=E0)
        tc = time.perf_counter() - t0
        cpu = {"value": X * Y * args.theories_cpu_steps / tc, "unit": "grid-point updates/s", "cores": 1,
               "kind": "port", "sample": f"NumPy oracle, 128x128, {args.theories_cpu_steps} steps, {tc:.2f} s"}
        print(f"[bench] {name} NumPy oracle: {cpu['value']:.3e} updates/s", file=sys.stderr, flush=True)
        out[name] = {"forms": forms, "cpu_baseline": cpu}
    print(json.dumps({"metric": "grid-pt updates/s + %HBM peak per theory",
                      "unit": "grid-point updates/s", "n_gpus": 1, "data": "synthetic", "steps": steps,
                      "warmup": warmup,
                      "config": {"workload": f"each of the nine theories on {n}x{n}, its own parameters and "
                                             "initial condition (bench.py THEORY_CASES)", "lattice": [n, n]},
                      "theories": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--graph-steps", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=1024, help="side of the CPU-baseline sample lattice")
    ap.add_argument("--cpu-steps", type=int, default=150, help="steps of the CPU-baseline sample (~10 s on 16 cores)")
    ap.add_argument("--numpy-steps", default="200,40",
                    help="first steps of the single-thread NumPy oracle on the Abelian-Higgs 128^2 and chiral 320^2 "
                         "workloads (0 skips one)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile shape of every pass")
    ap.add_argument("--tiles", default="", help="tile variants of passes 1-4 for the FP32 form, e.g. 1,1,1,2 "
                                                "(no autotuning)")
    ap.add_argument("--fused23", type=int, default=-1, help="with --tiles: 1 runs passes 2+3 as one sweep")
    ap.add_argument("--fused41", type=int, default=-1,
                    help="with --tiles: 1 runs pass 4 and the next step's pass 1 as one sweep")
    ap.add_argument("--pairs", type=int, default=-1,
                    help="with --tiles: 1 runs a step as the two paired sweeps, passes 1+2 and 3+4")
    ap.add_argument("--profile-only", action="store_true", help="FP32 form only, no CPU leg (for rocprof)")
    ap.add_argument("--reduced-form", choices=["fp32", "fp32c"], default="fp32c",
                    help="throughput workload: the FP32 form whose rate is the headline value (fp32c, the "
                         "default: with the compensated field update, the form that converges; fp32: plain)")
    ap.add_argument("--ttc-n", type=int, default=4096,
                    help="throughput workload: lattice side of the time-to-converge run (one anti-skyrmion)")
    ap.add_argument("--diff-xsize", type=float, default=100.0,
                    help="differentiation workload: box side of the 512^2 lattice")
    ap.add_argument("--diff-dt", type=float, default=0.3,
                    help="differentiation workload: time step in units of the spacing")
    ap.add_argument("--diff-max-steps", type=int, default=400000,
                    help="differentiation workload: step budget of each relaxation (the first one's tail, from "
                         "3e-6 to 1e-6, is slow: about 210 000 steps in all)")
    ap.add_argument("--theories-n", type=int, default=2048, help="theories workload: lattice side")
    ap.add_argument("--theories-cpu-steps", type=int, default=20,
                    help="theories workload: steps of the NumPy oracle on each theory's 128^2 sample")
    ap.add_argument("--workload", choices=["throughput", "precision", "batch", "converge", "differentiation",
                                           "theories"],
                    default="throughput",
                    help="precision: time-to-converge of the FP32 form + FP64 polish against the FP64 form; "
                         "batch: many baby-Skyrme relaxations of uneven sizes; converge: time-to-converge of the "
                         "Abelian-Higgs 128^2 and chiral 320^2 FP64 workloads; differentiation: dE*/dbeta by "
                         "backward at a relaxed 512^2 anisotropic minimum")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--nsk", type=int, default=1, help="precision workload: seeded anti-skyrmions")
    ap.add_argument("--demag", type=int, default=0, help="precision workload: 1 turns the demagnetising field on")
    ap.add_argument("--chunk", type=int, default=10000, help="precision workload: FP32 steps between polish checks")
    ap.add_argument("--polish", type=int, default=200, help="precision workload: FP64 steps per polish check")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-workload step rate of the default line")
    ap.add_argument("--no-precision", action="store_true",
                    help="throughput workload: skip the time-to-converge run of the FP32C form")
    ap.add_argument("--converge-cap", type=int, default=0,
                    help="converge workload: stop each case after at most this many steps (0: run to the tolerance)")
    ap.add_argument("--start", choices=["fp64", "reduced"], default="reduced",
                    help="precision workload: begin in FP64 and hand over to the FP32 form once FP64 converges "
                         "slower than it would (fp64), or begin in the FP32 form (reduced)")
    ap.add_argument("--lead-chunk", type=int, default=2000, help="precision workload: FP64 lead steps per check")
    ap.add_argument("--lead-rate", type=float, default=1.33,
                    help="precision workload: hand over to the FP32 form once an FP64 lead chunk shrinks max|EG| "
                         "by less than this factor")
    ap.add_argument("--form", choices=["fp32c", "fp32"], default="fp32c",
                    help="precision workload: FP32 phase with the compensated field update (fp32c) or plain (fp32)")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--tol", type=float, default=None,
                    help="max |EG| to reach (precision and batch workloads: 1e-6)")
    ap.add_argument("--max-steps", type=int, default=200000)
    ap.add_argument("--fp32-steps", type=int, default=20000,
                    help="batch workload: most FP32 steps before the FP64 retire phase")
    ap.add_argument("--batch-max-steps", type=int, default=60000,
                    help="batch workload: most steps in all (FP32 + FP64)")
    ap.add_argument("--check-every", type=int, default=500)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that measure the dominant kernel's HBM bytes")
    ap.add_argument("--traffic-child", type=int, default=0, metavar="K",
                    help="only run K eager steps of the chosen form and shapes from the seeded field and print "
                         "which of them ran arrested (the child that rocprofv3 --pmc counts bytes in)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates rocprof byte counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                          "launches": 3}), flush=True)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the relaxation core is single-GPU "
              "(multi-GPU halo exchange is a design note, not built).", file=sys.stderr)
        sys.exit(2)

    import torch
    import soliton_solver_amd as ssa
    from soliton_solver_amd import _lib
    if _lib.device_count() < 1:
        print("no GPU visible to HIP", file=sys.stderr)
        sys.exit(1)
    torch.cuda.init()

    if args.workload == "precision":
        args.tol = 1e-6 if args.tol is None else args.tol
        precision_workload(args)
        return
    if args.workload == "batch":
        args.tol = 1e-6 if args.tol is None else args.tol
        batch_workload(args)
        return
    if args.workload == "converge":
        converge_workload(args)
        return
    if args.workload == "theories":
        theories_workload(args)
        return
    if args.workload == "differentiation":
        args.tol = 1e-6 if args.tol is None else args.tol
        differentiation_workload(args)
        return

    n = args.n
    th, p, cfg = chiral_workload(n)
    print(f"[bench] building {n}x{n} chiral workload", file=sys.stderr, flush=True)
    form = args.reduced_form
    comp = form == "fp32c"
    nf, w, N = 4, 4, n * n
    tiles_arg = [int(v) for v in args.tiles.split(",")] if args.tiles else None
    if args.traffic_child:
        traffic_child(th, p, cfg, form, tiles_arg, args)
        return
    sim32, E32, err32, ea32, t32 = time_form(th, p, cfg, form, args.steps, args.warmup, args.graph_steps,
                                             autotune=not args.no_autotune, tiles=tiles_arg,
                                             fused23=args.fused23, fused41=args.fused41, pairs=args.pairs)
    tiles32 = sim32.tiles
    print(f"[bench] {form} tile rows per pass {tiles32['rows']}", file=sys.stderr, flush=True)
    if "ms" in tiles32:
        for k, row in enumerate(tiles32["ms"]):
            print(f"[bench]   pass {k + 1} ms by shape variant: " + " ".join(f"{x:.4f}" for x in row),
                  file=sys.stderr, flush=True)
        for key, label in (("ms_fused23", "fused passes 2+3"), ("ms_fused41", "fused passes 4+1"),
                           ("ms_pairs12", "paired sweep 1+2"), ("ms_pairs34", "paired sweep 3+4")):
            print(f"[bench]   {label} ms by shape variant: " + " ".join(f"{x:.4f}" for x in tiles32[key]),
                  file=sys.stderr, flush=True)
        print(f"[bench]   chosen: fused23 {tiles32['fused23']}, fused41 {tiles32['fused41']}, "
              f"paired {tiles32['pairs']}", file=sys.stderr, flush=True)
    ups32 = N * args.steps / t32
    print(f"[bench] {form} form: {t32 / args.steps * 1e3:.3f} ms/step, {ups32:.3e} updates/s", file=sys.stderr,
          flush=True)

    # dominant-kernel roofline from HIP events on the context stream, over the
    # timed steps: a second context with the same seeded field, shapes and
    # warm-up reaches the state they started from and replays them eagerly,
    # pass by pass (the same trajectory bit for bit); its algorithmic bytes per
    # launch follow the arrest decisions of those steps (an arrested step reads
    # no V)
    timed_arrested = sim32.timed_arrested
    sim32.close()
    simr = ssa.Simulation(p, th, precision=form)
    simr.initialize(cfg)
    simr.set_tile_variants(tiles32["variants"])
    simr.set_fused_stages(bool(tiles32.get("fused23", False)), bool(tiles32["fused41"]), pairs=bool(tiles32["pairs"]))
    if args.graph_steps:
        simr.use_graphs(args.graph_steps)
    E_start = simr.compute_energy()
    E_start, _, _ = simr.steps(E_start, args.warmup)
    nroof = args.steps
    passes = simr.time_passes(E_start, nroof)
    E_replay = simr.history(nroof)[0]
    simr.close()
    kills = arrested_steps(E_replay, E_start)
    if E_replay[-1] != E32 or int(kills.sum()) != timed_arrested:
        print(f"[bench] roofline replay: last energy {E_replay[-1]!r} against {E32!r}, {int(kills.sum())} "
              f"arrests against {timed_arrested}", file=sys.stderr, flush=True)
    sched = schedule(tiles32, comp)
    words_dom = launch_words(sched, kills)
    alg_dom = words_dom * nf * w * N
    dom_gbs = alg_dom / (passes[4] * 1e-3) / 1e9
    step_ms_kernels = float(passes.sum())
    alg_step = sched["step_words"] * nf * w * N
    step_gbs = alg_step / (step_ms_kernels * 1e-3) / 1e9

    result64 = None
    if not args.profile_only:
        sim64, E64, err64, ea64, t64 = time_form(th, p, cfg, "fp64", args.steps, args.warmup, args.graph_steps,
                                                 autotune=not args.no_autotune)
        tiles64 = sim64.tiles
        ups64 = N * args.steps / t64
        relE = abs(E32 - E64) / abs(E64)
        sim64.close()
        s64 = schedule(tiles64, False)
        result64 = {"updates_per_s": ups64, "ms_per_step": t64 / args.steps * 1e3,
                    "timed_steps_arrested": sim64.timed_arrested,
                    "hbm_frac_alg": ups64 * 16 * nf * 8 / (HBM_PEAK_GBS * 1e9),
                    "hbm_frac_schedule": ups64 * s64["step_words"] * nf * 8 / (HBM_PEAK_GBS * 1e9),
                    "words_per_update": s64["step_words"],
                    "energy": E64, f"{form}_vs_fp64_relE_after_steps": relE,
                    "relE_note": "both forms from the same field after the same timed steps, mid-relaxation; "
                                 "relaxed energies: time_to_converge",
                    "tile_rows_per_pass": tiles64["rows"]}
        print(f"[bench] fp64 form: {t64 / args.steps * 1e3:.3f} ms/step, {ups64:.3e} updates/s, "
              f"relE({form} vs fp64)={relE:.2e}", file=sys.stderr, flush=True)

    other = None
    if not args.profile_only:
        oform = "fp32" if comp else "fp32c"
        simo, Eo, _, _, to = time_form(th, p, cfg, oform, args.steps, args.warmup, args.graph_steps,
                                       autotune=not args.no_autotune)
        upso = N * args.steps / to
        so = schedule(simo.tiles, oform == "fp32c")
        other = {"form": oform, "updates_per_s": upso, "ms_per_step": to / args.steps * 1e3,
                 "hbm_frac_alg": upso * 16 * nf * 4 / (HBM_PEAK_GBS * 1e9),
                 "hbm_frac_schedule": upso * so["step_words"] * nf * 4 / (HBM_PEAK_GBS * 1e9),
                 "words_per_update": so["step_words"],
                 "speedup_vs_fp64": (upso / result64["updates_per_s"]) if result64 else None,
                 "tile_rows_per_pass": simo.tiles["rows"],
                 "note": ("plain FP32 storage: its steps round away increments below an FP32 ulp, so alone it "
                          "stalls above the FP64 minimum and does not reach Delta E / E <= 1e-7 without a long "
                          "FP64 tail" if oform == "fp32" else
                          "compensated field update (FP32 hi + lo words), the form that converges")}
        simo.close()
        print(f"[bench] {oform} form: {to / args.steps * 1e3:.3f} ms/step, {upso:.3e} updates/s",
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu and not args.profile_only:
        cpu = cpu_baseline(args, n)

    batch = None
    if not args.no_batch and not args.profile_only:
        batch = batch_throughput(args)
        print(f"[bench] batch: {batch['ms_per_step']:.3f} ms per step, {batch['updates_per_s']:.3e} updates/s",
              file=sys.stderr, flush=True)

    # time to converge of the compensated FP32 form against the FP64 form, to
    # max|EG| < 1e-6 from one seeded field: one anti-skyrmion at 4096^2
    # (demagnetisation off; with it on, the saddle field psi keeps the flow
    # killing its velocity nearly every step and no form converges in a bounded run)
    ttc = None
    if not args.no_precision and not args.profile_only:
        import copy
        pa = copy.copy(args)
        pa.n, pa.nsk, pa.demag, pa.tol, pa.form, pa.seed = args.ttc_n, 1, 0, 1e-6, "fp32c", 4
        pr = precision_workload(pa, emit=False)
        ttc = {"workload": pr["config"]["workload"], "speedup_vs_fp64": pr["value"], "converged": pr["converged"],
               "fp64_seconds": pr["fp64_form"]["seconds"], "fp64_steps": pr["fp64_form"]["steps"],
               "fp32c_seconds": pr["fp32_phase"]["seconds"] + pr["fp64_polish"]["seconds"],
               "fp32c_steps": pr["fp32_phase"]["steps"], "fp64_polish_steps": pr["fp64_polish"]["steps_total"],
               "relE_vs_fp64_form": pr["relE_vs_fp64_form"], "max_abs_dm": pr["max_dm_vs_fp64_form"],
               "skyrmion_number": pr["skyrmion_number"],
               "skyrmion_number_identical": pr["skyrmion_number"]["identical_rounded"]}
        print(f"[bench] time to converge ({args.ttc_n}^2, fp32c vs fp64): {ttc['speedup_vs_fp64']:.2f}x, "
              f"relE {ttc['relE_vs_fp64_form']:.1e}", file=sys.stderr, flush=True)

    kernel = sched["kernel"]
    traffic = None
    if not args.no_traffic and not args.profile_only:
        child = ["--tiles", ",".join(str(v) for v in tiles32["variants"]),
                 "--fused23", str(int(tiles32.get("fused23", False))), "--fused41", str(int(tiles32["fused41"])),
                 "--pairs", str(int(tiles32["pairs"])), "--n", str(n),
                 "--reduced-form", form]
        traffic = measure_traffic(child, kernel)
        if traffic:
            ka = np.asarray(traffic["arrested"], dtype=bool)
            traffic["alg_bytes"] = launch_words(sched, ka) * nf * w * N
            traffic["note"] = (f"rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, over the {len(ka)} eager steps of a "
                               f"child run from the seeded field ({int(ka.sum())} arrested), copy-calibrated; "
                               f"{traffic['bytes'] / traffic['alg_bytes']:.2f}x the algorithmic bytes of those "
                               "launches")
            print(f"[bench] traffic of {kernel}: {traffic['bytes'] / 1e9:.3f} GB per launch "
                  f"({traffic['bytes'] / traffic['alg_bytes']:.2f}x the algorithmic bytes of the same launches, "
                  f"{int(ka.sum())} of {len(ka)} arrested)", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": ups32,
        "unit": "grid-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "form": form,
        "timed_steps_arrested": timed_arrested,
        "tile_variants": tiles32["variants"],
        "fused23": int(tiles32.get("fused23", False)),
        "fused41": int(tiles32["fused41"]),
        "paired_sweeps": int(tiles32["pairs"]),
        "config": {"workload": f"chiral magnet {n}x{n}, h=10/319, Heusler DMI, demag, 8 seeded anti-skyrmions; "
                               "FP32 fields+stencils, FP64 energy sums and arrest test"
                               + ("; compensated field update (FP32 hi + lo)" if comp else ""),
                   "lattice": [n, n], "fields": nf, "graph_steps": args.graph_steps,
                   "tile_rows_per_pass": tiles32["rows"], "tile_shape": "64 y x rows x",
                   "schedule": sched["name"], "words_per_update": sched["step_words"],
                   "parallelism": "single GPU"},
        "hbm_frac_alg": ups32 * 16 * nf * w / (HBM_PEAK_GBS * 1e9),
        "hbm_frac_alg_note": "updates/s x 16 nf w bytes (the fixed algorithmic bytes of one update) / 8 TB/s",
        "hbm_frac_schedule": ups32 * sched["step_words"] * nf * w / (HBM_PEAK_GBS * 1e9),
        "roofline": {"bound": "hbm", "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom_gbs / HBM_PEAK_GBS, "traffic": traffic["bytes"] if traffic else None,
                     "traffic_read_write": [traffic["read_bytes"], traffic["write_bytes"]] if traffic else None,
                     "traffic_alg_bytes_per_launch": traffic["alg_bytes"] if traffic else None,
                     "traffic_note": traffic["note"] if traffic else None,
                     "kernel": kernel, "alg_words_per_site": words_dom * nf,
                     "alg_words_per_site_note": f"{sched['dominant_words'] * nf} per site in a step with the velocity "
                                                f"free, {sched['arrested_words'] * nf} in an arrested one (no V read); "
                                                f"{int(kills.sum())} of the {nroof} timed steps arrested",
                     "launches_note": f"the {nroof} timed steps replayed in a second context from the same seeded "
                                      "field, shapes and warm-up, eagerly, each pass between HIP events on the "
                                      "context's stream; the replay "
                                      + ("ends on the timed run's energy bit for bit"
                                         if E_replay[-1] == E32 else "does NOT end on the timed run's energy"),
                     "alg_bytes_per_launch": alg_dom, "avg_launch_ms": passes[4],
                     "step_kernels_ms": step_ms_kernels, "step_kernels_gbs": step_gbs,
                     "pass_ms": [float(x) for x in passes]},
        "fp64_form": result64,
        "speedup_vs_fp64": (ups32 / result64["updates_per_s"]) if result64 else None,
        ("fp32_form" if comp else "fp32c_form"): other,
        "time_to_converge": ttc,
        "batch": batch,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
