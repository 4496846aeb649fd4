"""Benchmark of the relaxation core on one MI355X.

Workload (default): chiral magnet, 4096 x 4096 lattice at fixed spacing
h = 10/319 (xsize = 4095 h), Heusler DMI, demagnetisation on, unit
magnetisation, coupling constants J, K, D, M, B = 40e-12, 0.8e6, 4e-3, 580e3, 0;
initial field = 8 seeded anti-skyrmions (centres uniform in the central 80 %,
radii uniform in [0.3, 0.6]) plus seeded Gaussian noise sigma = 1e-3,
renormalised.  One step = one arrested-Newton RK4 step over every site.

The headline ``value`` is grid-point updates/s of the FP32-stencil /
FP64-reduction form; the plain FP64 form is timed beside it on the same
initial field.  ``roofline`` prices the dominant kernel (the RK4 pass 4
launch, 7 nf w bytes per site) from HIP events on the context's stream.
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


def chiral_workload(n: int, seed: int = 0, nsk: int = 8):
    import soliton_solver_amd as ssa
    th = ssa.load_theory("Chiral magnet")
    h = 10.0 / 319.0
    L = h * (n - 1)
    p = th.params.default_params(xlen=n, ylen=n, xsize=L, ysize=L, J=40e-12, K=0.8e6, D=4e-3, M=580e3, B=0.0,
                                 dmi_term="Heusler", demag=True, unit_magnetization=True)
    rng = np.random.default_rng(seed)
    sk = [(rng.uniform(0.1 * L, 0.9 * L), rng.uniform(0.1 * L, 0.9 * L), rng.uniform(0.3, 0.6))
          for _ in range(nsk)]
    return th, p, {"mode": "skyrmions", "ansatz": "anti", "skyrmions": sk, "noise": 1e-3, "seed": seed + 1}


def time_form(th, p, cfg, precision, steps, warmup, graph_steps, F0=None, autotune=True, tiles=None, fused23=-1):
    import soliton_solver_amd as ssa
    sim = ssa.Simulation(p, th, precision=precision)
    if F0 is None:
        sim.initialize(cfg)
    else:
        sim.initialize({"mode": "array", "field": F0})
    if tiles is not None:
        sim.set_tile_variants(tiles)
        if fused23 >= 0:
            sim.set_fused_stages(bool(fused23))
        sim.tiles = sim.tile_variants()
    else:
        sim.tiles = sim.autotune_tiles(reps=5) if autotune else sim.tile_variants()
    if graph_steps:
        sim.use_graphs(graph_steps)
    E = sim.compute_energy()
    E, _, _ = sim.steps(E, warmup)
    sim.synchronize()
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E, err, ea = sim.steps(E, steps)
    sim.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return sim, E, err, ea, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--graph-steps", type=int, default=20)
    ap.add_argument("--cpu-sample", type=int, default=1024, help="side of the CPU-baseline sample lattice")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tile shape of every pass")
    ap.add_argument("--tiles", default="", help="tile variants of passes 1-4 for the FP32 form, e.g. 1,1,1,2 "
                                                "(no autotuning)")
    ap.add_argument("--fused23", type=int, default=-1, help="with --tiles: 1 runs passes 2+3 as one sweep")
    ap.add_argument("--profile-only", action="store_true", help="FP32 form only, no CPU leg (for rocprof)")
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

    n = args.n
    th, p, cfg = chiral_workload(n)
    print(f"[bench] building {n}x{n} chiral workload", file=sys.stderr, flush=True)
    sim32, E32, err32, ea32, t32 = time_form(th, p, cfg, "fp32", args.steps, args.warmup, args.graph_steps,
                                             autotune=not args.no_autotune,
                                             tiles=[int(v) for v in args.tiles.split(",")] if args.tiles else None,
                                             fused23=args.fused23)
    tiles32 = sim32.tiles
    print(f"[bench] fp32 tile rows per pass {tiles32['rows']}", file=sys.stderr, flush=True)
    if "ms" in tiles32:
        for k, row in enumerate(tiles32["ms"]):
            print(f"[bench]   pass {k + 1} ms by shape variant: " + " ".join(f"{x:.4f}" for x in row),
                  file=sys.stderr, flush=True)
        print(f"[bench]   fused passes 2+3 ms by strip variant: "
              + " ".join(f"{x:.4f}" for x in tiles32["ms_fused23"]) + f"; fused: {tiles32['fused23']}",
              file=sys.stderr, flush=True)
    F0 = None
    N = n * n
    nf = 4
    ups32 = N * args.steps / t32
    print(f"[bench] fp32 form: {t32 / args.steps * 1e3:.3f} ms/step, {ups32:.3e} updates/s", file=sys.stderr,
          flush=True)

    # dominant-kernel roofline from HIP events on the context stream
    passes = sim32.time_passes(E32, 10)
    w = 4
    alg_pass4 = 7 * nf * w * N
    pass4_gbs = alg_pass4 / (passes[4] * 1e-3) / 1e9
    step_ms_kernels = float(passes.sum())
    alg_step = 16 * nf * w * N
    step_gbs = alg_step / (step_ms_kernels * 1e-3) / 1e9
    sim32.close()

    result64 = None
    if not args.profile_only:
        sim64, E64, err64, ea64, t64 = time_form(th, p, cfg, "fp64", args.steps, args.warmup, args.graph_steps,
                                                 autotune=not args.no_autotune)
        tiles64 = sim64.tiles
        ups64 = N * args.steps / t64
        relE = abs(E32 - E64) / abs(E64)
        sim64.close()
        result64 = {"updates_per_s": ups64, "ms_per_step": t64 / args.steps * 1e3,
                    "hbm_frac_alg": ups64 * 16 * nf * 8 / (HBM_PEAK_GBS * 1e9),
                    "energy": E64, "fp32_vs_fp64_relE_after_steps": relE,
                    "tile_rows_per_pass": tiles64["rows"]}
        print(f"[bench] fp64 form: {t64 / args.steps * 1e3:.3f} ms/step, {ups64:.3e} updates/s, "
              f"relE(fp32 vs fp64)={relE:.2e}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu and not args.profile_only:
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
        threads = max(1, min(threads, 16))
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
                         f"host {model or platform.processor()}, os.cpu_count()={os.cpu_count()}"}

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
        "tile_variants": tiles32["variants"],
        "fused23": int(tiles32.get("fused23", False)),
        "config": {"workload": f"chiral magnet {n}x{n}, h=10/319, Heusler DMI, demag, 8 seeded anti-skyrmions; "
                               "FP32 fields+stencils, FP64 energy sums and arrest test",
                   "lattice": [n, n], "fields": nf, "graph_steps": args.graph_steps,
                   "tile_rows_per_pass": tiles32["rows"], "tile_shape": "64 y x rows x",
                   "fused_passes_2_3": tiles32.get("fused23", False),
                   "parallelism": "single GPU"},
        "hbm_frac_alg": ups32 * 16 * nf * w / (HBM_PEAK_GBS * 1e9),
        "roofline": {"bound": "hbm", "achieved": pass4_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": pass4_gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"stage_pass<ChiralMagnet, float, 4, {tiles32['rows'][3]}>",
                     "alg_bytes_per_launch": alg_pass4, "avg_launch_ms": passes[4],
                     "step_kernels_gbs": step_gbs, "pass_ms": [float(x) for x in passes]},
        "fp64_form": result64,
        "speedup_fp32_vs_fp64": (ups32 / result64["updates_per_s"]) if result64 else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
