"""Launch-shape limits of the streaming form (GPU).

The streaming passes read each array of nf planes through one buffer resource
with 32-bit byte offsets below 2^31 (a site outside the lattice gets an offset
beyond the array and reads 0).  A lattice whose arrays reach 2^31 bytes runs
the tile form whatever shape variant is asked for.  Chiral magnet, FP64,
8192 x 8193 sites: 4 planes x 8 B x N = 2 147 745 792 bytes, just past the
limit.  Asking for streaming strips (one and two wavefronts wide) must give
fields bit-identical to asking for the default LDS tiles, two steps from the
same device-initialised field; the sweeps that exist only as strips (the
paired 1+2 / 3+4 and, at this size, the fused 2+3) are refused with an
error.
"""
from __future__ import annotations

import numpy as np
import pytest

import soliton_solver_amd as ssa

pytestmark = pytest.mark.gpu


def _run(sim, variants):
    sim.set_tile_variants(variants)
    sim.initialize({"mode": "initial", "ansatz": "anti", "noise": 1e-3, "seed": 3, "device": True})
    E = sim.compute_energy()
    E, err, _ = sim.steps(E, 2)
    return E, sim.get_field()


def test_strips_fall_back_to_tiles_beyond_31bit_arrays():
    th = ssa.load_theory("Chiral magnet")
    X, Y = 8192, 8193
    h = 10.0 / 319.0
    assert 4 * X * Y * 8 >= 2 ** 31
    p = th.params.default_params(xlen=X, ylen=Y, xsize=h * (X - 1), ysize=h * (Y - 1), dmi_term="Heusler",
                                 demag=True, unit_magnetization=True)
    sim = ssa.Simulation(p, th)
    try:
        E0, F0 = _run(sim, [0, 0, 0, 0])
        assert np.isfinite(E0)
        for variants in ([3, 3, 3, 3], [6, 6, 6, 6]):
            E1, F1 = _run(sim, variants)
            print(f"variants {variants}: E {E1!r} (tiles {E0!r})", flush=True)
            assert E1 == E0
            assert np.array_equal(F1, F0)
        with pytest.raises(ValueError, match="paired sweeps"):
            sim.set_fused_stages(False, False, pairs=True)
        with pytest.raises(ValueError, match="cannot share one sweep"):
            sim.set_fused_stages(True, False)
    finally:
        sim.close()
