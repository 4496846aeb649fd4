"""Native kernel handles of the Ginzburg-Landau superconductor (csrc/theory_gl.h,
restating the reference's theories/ginzburg_landau_superconductor/kernels.py)."""
from ..._lib import THEORY_GINZBURG_LANDAU
from ...core.integrator import NativeKernel, make_do_gradient_step_kernel

THEORY_ID = THEORY_GINZBURG_LANDAU
CHARGE_KINDS = {"vortex_number": 0}

do_gradient_step_kernel = make_do_gradient_step_kernel(THEORY_ID)
compute_energy_kernel = NativeKernel(THEORY_ID, "energy")
compute_vortex_number_kernel = NativeKernel(THEORY_ID, "charge:vortex_number")
