"""Text output bundle and resume (the on-disk format of the reference's
core/io.py:1-241 and theories/*/io.py).

Same file layout so the reference's plotting scripts read it unchanged:
every value printed with ``{:.<precision>g}`` and followed by a tab; a density
file is one line per x (y along the line); ``Field.dat`` is every component
in turn, each block followed by an empty line; ``LatticePoints.dat`` holds
xlen ylen and ``LatticeVectors.dat`` x_max 0 0 y_max (x_max, y_max = the grid
coordinates of the last site).  Rows are formatted with NumPy's C loop
(``'%.32g' % v`` and ``'{:.32g}'.format(v)`` give the same text for every
double), which keeps 4096^2 fields to seconds.

``read_field_dat`` parses a ``Field.dat`` back into the flat reference layout
for resuming a relaxation.
"""
from __future__ import annotations

import os

import numpy as np


def _flat(a: int, x: int, y: int, xlen: int, ylen: int) -> int:
    return y + x * ylen + a * xlen * ylen


def _write_block(f, block: np.ndarray, precision: int) -> None:
    # each value followed by a tab, each row by a newline (core/io.py:55-70)
    np.savetxt(f, np.asarray(block, dtype=np.float64).reshape(block.shape[0], -1), fmt=f"%.{precision}g",
               delimiter="\t", newline="\t\n")


def output_field_dat(field_flat: np.ndarray, path: str, xlen: int, ylen: int, nfields: int,
                     precision: int = 32) -> None:
    """core/io.py:44-70: every component as an xlen x ylen block, blocks separated by an empty line."""
    F = np.asarray(field_flat, dtype=np.float64).reshape(-1)[: nfields * xlen * ylen].reshape(nfields, xlen, ylen)
    with open(path, "w", newline="\n") as f:
        for a in range(nfields):
            _write_block(f, F[a], precision)
            f.write("\n")


def output_iteration_data_dat(values, path: str, precision: int = 32) -> None:
    """core/io.py:72-90: one line of tab-terminated values."""
    fmt = f"{{:.{precision}g}}"
    with open(path, "w", newline="\n") as f:
        for v in values:
            f.write(fmt.format(float(v)))
            f.write("\t")


def output_density_data_dat(density_flat: np.ndarray, a: int, path: str, xlen: int, ylen: int,
                            precision: int = 32) -> None:
    """core/io.py:92-118: component ``a`` of a flat a-major array as one xlen x ylen block."""
    D = np.asarray(density_flat, dtype=np.float64).reshape(-1)
    block = D[a * xlen * ylen:(a + 1) * xlen * ylen].reshape(xlen, ylen)
    with open(path, "w", newline="\n") as f:
        _write_block(f, block, precision)


def output_data_bundle_core(output_dir: str, *, h_Field: np.ndarray, h_grid: np.ndarray, xlen: int, ylen: int,
                            number_coordinates: int, number_total_fields: int, precision: int,
                            bundle_spec: list[tuple[str, int, str]], arrays: dict[str, np.ndarray],
                            lattice_points_name: str = "LatticePoints.dat",
                            lattice_vectors_name: str = "LatticeVectors.dat",
                            field_dump_name: str = "d_Field.dat") -> None:
    """core/io.py:120-241: lattice files, the full field dump, then each (array, component, file) of the spec."""
    os.makedirs(output_dir, exist_ok=True)
    g = np.asarray(h_grid, dtype=np.float64).reshape(-1)
    x_max = float(g[_flat(0, xlen - 1, ylen - 1, xlen, ylen)]) if number_coordinates >= 1 else 0.0
    y_max = float(g[_flat(1, xlen - 1, ylen - 1, xlen, ylen)]) if number_coordinates >= 2 else 0.0
    points = np.array([xlen, ylen], dtype=np.float64)
    vectors = np.array([x_max, 0.0, 0.0, y_max], dtype=np.float64)
    output_iteration_data_dat(points[:number_coordinates], os.path.join(output_dir, lattice_points_name), precision)
    output_iteration_data_dat(vectors[:number_coordinates * number_coordinates],
                              os.path.join(output_dir, lattice_vectors_name), precision)
    output_field_dat(h_Field, os.path.join(output_dir, field_dump_name), xlen, ylen, number_total_fields, precision)
    for key, a, filename in bundle_spec:
        arr = arrays.get(key)
        if arr is None:
            continue
        output_density_data_dat(arr, int(a), os.path.join(output_dir, filename), xlen, ylen, precision)


def read_field_dat(path: str, xlen: int, ylen: int, nfields: int) -> np.ndarray:
    """Parse a Field.dat written by ``output_field_dat`` into shape (nfields, xlen, ylen)."""
    with open(path) as f:
        vals = np.array(f.read().split(), dtype=np.float64)
    if vals.size != nfields * xlen * ylen:
        raise ValueError(f"{path}: {vals.size} values, expected {nfields}*{xlen}*{ylen}")
    return vals.reshape(nfields, xlen, ylen)
