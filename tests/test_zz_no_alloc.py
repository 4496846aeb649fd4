"""ss_energy_dev_noalloc (the allocation-free device-pointer energy, partials
in the library's static device scratch) against ss_energy_dev (partials in
the context's allocation): the same kernels and summation order, so the two
totals are bit-identical.  Two calls in a row on one GPU, the second on a
smaller lattice (fewer partials in the shared scratch array), then the first
lattice again."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import soliton_solver_amd as ssa
from soliton_solver_amd import _lib
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _energies(sim, F, torch):
    L = _lib.lib()
    x = torch.tensor(np.ascontiguousarray(F).ravel(), dtype=torch.float64, device="cuda")
    e_ref = torch.zeros(1, dtype=torch.float64, device="cuda")
    e_na = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    stream = C.c_void_p(int(torch.cuda.current_stream().cuda_stream))
    _lib.check(L.ss_energy_dev(sim._ctx, C.c_void_p(x.data_ptr()), C.c_void_p(e_ref.data_ptr()), stream), sim._ctx)
    _lib.check(L.ss_energy_dev_noalloc(sim._ctx, C.c_void_p(x.data_ptr()), C.c_void_p(e_na.data_ptr()), stream),
               sim._ctx)
    torch.cuda.synchronize()
    return float(e_ref.item()), float(e_na.item())


def test_noalloc_energy_matches_energy_dev_on_two_calls():
    torch = pytest.importorskip("torch")
    th = ssa.load_theory("Chiral magnet")
    sims = []
    try:
        results = []
        for n, m, seed in ((160, 144, 1), (40, 36, 2), (160, 144, 3)):
            p = H.chiral_params(th, n, m)
            sim = ssa.Simulation(p, th)
            sims.append(sim)
            p_i, p_f = H.packed(th, p)
            F = H.chiral_field(th, p_i, p_f, seed=seed)
            e_ref, e_na = _energies(sim, F, torch)
            print(f"[no-alloc] {n}x{m}: energy_dev {e_ref!r} noalloc {e_na!r}", flush=True)
            assert e_na == e_ref
            assert np.isfinite(e_ref)
            results.append(e_ref)
        assert results[0] != results[1]
    finally:
        for s in sims:
            s.close()


def test_noalloc_energy_refuses_non_fp64_contexts():
    torch = pytest.importorskip("torch")
    th = ssa.load_theory("Chiral magnet")
    sim = ssa.Simulation(H.chiral_params(th, 48, 40), th, precision="fp32")
    try:
        x = torch.zeros(4 * 48 * 40, dtype=torch.float64, device="cuda")
        e = torch.zeros(1, dtype=torch.float64, device="cuda")
        rc = _lib.lib().ss_energy_dev_noalloc(sim._ctx, C.c_void_p(x.data_ptr()), C.c_void_p(e.data_ptr()), None)
        assert rc == -1 and "FP64" in _lib.error_text(sim._ctx)
    finally:
        sim.close()
