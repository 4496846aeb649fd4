"""Native kernel handles of the baby Skyrme model (csrc/theory_bs.h, restating the
reference's theories/baby_skyrme/kernels.py)."""
from ..._lib import THEORY_BABY_SKYRME
from ...core.integrator import NativeKernel, make_do_gradient_step_kernel, make_do_rk4_kernel

THEORY_ID = THEORY_BABY_SKYRME
CHARGE_KINDS = {"skyrmion_number": 0}

do_gradient_step_kernel = make_do_gradient_step_kernel(THEORY_ID)
do_rk4_kernel = make_do_rk4_kernel(THEORY_ID)
compute_energy_kernel = NativeKernel(THEORY_ID, "energy")
compute_skyrmion_number_kernel = NativeKernel(THEORY_ID, "charge:skyrmion_number")
