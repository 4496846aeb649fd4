"""Output bundle of this theory (the file set of the reference's
theories/chiral_magnet/io.py, written by ``core.io.output_data_bundle_core``)."""
from __future__ import annotations

import numpy as np

from ...core.io import output_data_bundle_core

# (array, component, file) in the reference's order: the on-disk format.
BUNDLE_SPEC = [
    ("grid", 0, "xGrid.dat"),
    ("grid", 1, "yGrid.dat"),
    ("Field", 0, "MagnetField1.dat"),
    ("Field", 1, "MagnetField2.dat"),
    ("Field", 2, "MagnetField3.dat"),
    ("BaryonDensity", 0, "SkyrmionChargeDensity.dat"),
    ("EnergyDensity", 0, "EnergyDensity.dat"),
    ("MagneticChargeDensity", 0, "MagneticChargeDensity.dat"),
]


def output_data_bundle(output_dir: str, h_Field: np.ndarray,
                       h_EnergyDensity: np.ndarray,
                       h_BaryonDensity: np.ndarray,
                       h_MagneticChargeDensity: np.ndarray,
                       h_grid: np.ndarray, xlen: int, ylen: int,
                       number_coordinates: int = 2, number_total_fields: int = 4, precision: int = 32) -> None:
    arrays = {
        "grid": h_grid,
        "Field": h_Field,
        "EnergyDensity": h_EnergyDensity,
        "BaryonDensity": h_BaryonDensity,
        "MagneticChargeDensity": h_MagneticChargeDensity,
    }
    output_data_bundle_core(output_dir, h_Field=h_Field, h_grid=h_grid, xlen=xlen, ylen=ylen,
                            number_coordinates=number_coordinates, number_total_fields=number_total_fields,
                            precision=precision, bundle_spec=BUNDLE_SPEC, arrays=arrays,
                            lattice_points_name="LatticePoints.dat", lattice_vectors_name="LatticeVectors.dat",
                            field_dump_name="Field.dat")
