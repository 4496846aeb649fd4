"""Output bundle text format and resume (host side, no GPU).

The writers must produce byte-identical files to the reference's value-by-value
loops (restated in oracle/io_format.py) for every precision the reference
offers, and Field.dat must read back to the same doubles at precision 32.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import io_format as oio
from soliton_solver_amd.core import io as sio


@pytest.mark.parametrize("precision", [32, 17, 8])
def test_writers_match_reference_text(tmp_path, precision):
    rng = np.random.default_rng(precision)
    X, Y, nf = 7, 5, 3
    F = rng.normal(size=nf * X * Y) * 10.0 ** rng.integers(-12, 12, size=nf * X * Y)
    F[3] = 0.0
    F[4] = -0.0
    p = os.path.join(tmp_path, "Field.dat")
    sio.output_field_dat(F, p, X, Y, nf, precision=precision)
    assert open(p).read() == oio.field_dat_text(F, X, Y, nf, precision)
    for a in range(nf):
        q = os.path.join(tmp_path, f"d{a}.dat")
        sio.output_density_data_dat(F, a, q, X, Y, precision=precision)
        assert open(q).read() == oio.density_dat_text(F, a, X, Y, precision)
    r = os.path.join(tmp_path, "v.dat")
    sio.output_iteration_data_dat([X, Y, 1.25], r, precision=precision)
    assert open(r).read() == oio.values_dat_text([X, Y, 1.25], precision)


def test_field_dat_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(1)
    F = rng.normal(size=(4, 9, 6))
    p = os.path.join(tmp_path, "Field.dat")
    sio.output_field_dat(F.ravel(), p, 9, 6, 4)
    assert np.array_equal(sio.read_field_dat(p, 9, 6, 4), F)
    with pytest.raises(ValueError):
        sio.read_field_dat(p, 9, 6, 5)


def test_bundle_core_files(tmp_path):
    X, Y = 4, 3
    g = np.empty((2, X, Y))
    g[0] = 0.5 * np.arange(X)[:, None]
    g[1] = 0.25 * np.arange(Y)[None, :]
    F = np.arange(2 * X * Y, dtype=np.float64)
    E = np.arange(X * Y, dtype=np.float64) * 0.1
    spec = [("grid", 0, "xGrid.dat"), ("grid", 1, "yGrid.dat"), ("EnergyDensity", 0, "EnergyDensity.dat"),
            ("Missing", 0, "Missing.dat")]
    sio.output_data_bundle_core(str(tmp_path), h_Field=F, h_grid=g.ravel(), xlen=X, ylen=Y, number_coordinates=2,
                                number_total_fields=2, precision=32, bundle_spec=spec,
                                arrays={"grid": g.ravel(), "EnergyDensity": E}, field_dump_name="Field.dat")
    names = sorted(os.listdir(tmp_path))
    assert names == ["EnergyDensity.dat", "Field.dat", "LatticePoints.dat", "LatticeVectors.dat", "xGrid.dat",
                     "yGrid.dat"]
    assert open(os.path.join(tmp_path, "LatticePoints.dat")).read() == "4\t3\t"
    assert open(os.path.join(tmp_path, "LatticeVectors.dat")).read() == "1.5\t0\t0\t0.5\t"
    assert open(os.path.join(tmp_path, "yGrid.dat")).read() == oio.density_dat_text(g.ravel(), 1, X, Y)


def test_oracle_text_is_the_correctly_rounded_decimal():
    """Each value the oracle writers emit is the double's exact binary value
    rounded to ``precision`` significant decimal digits (round half even), as
    exact Decimal arithmetic gives it, and parses back to the same double at 32."""
    from decimal import ROUND_HALF_EVEN, Context, Decimal
    values = [0.1, -1.0 / 3.0, 6.02214076e23, 5e-324, 1.0, 123456789.125, -2.5e-300, 2.0 ** 70 + 1.0]
    for x in values:
        for p in (32, 17, 8):
            text = oio.values_dat_text([x], p)
            assert text.endswith("\t") and text.count("\t") == 1
            assert Decimal(text[:-1]) == Context(prec=p, rounding=ROUND_HALF_EVEN).plus(Decimal(x))
        assert float(oio.values_dat_text([x], 32)[:-1]) == x
