"""Native kernel handles of the liquid crystal (csrc/theory_lc.h, restating the
reference's theories/liquid_crystal/kernels.py)."""
from ..._lib import THEORY_LIQUID_CRYSTAL
from ...core.integrator import NativeKernel, make_do_gradient_step_kernel, make_do_rk4_kernel

THEORY_ID = THEORY_LIQUID_CRYSTAL
CHARGE_KINDS = {"skyrmion_number": 0, "electric_charge": 1}

do_gradient_step_kernel = make_do_gradient_step_kernel(THEORY_ID)
do_rk4_kernel = make_do_rk4_kernel(THEORY_ID)   # unit-norm constraint + velocity projection
compute_energy_kernel = NativeKernel(THEORY_ID, "energy")
compute_skyrmion_number_kernel = NativeKernel(THEORY_ID, "charge:skyrmion_number")
compute_electric_charge_kernel = NativeKernel(THEORY_ID, "charge:electric_charge")
