"""Native kernel handles of the anisotropic superconductor (csrc/theory_as.h,
restating the reference's theories/anisotropic_superconductor/kernels.py)."""
from ..._lib import THEORY_ANISOTROPIC_SC
from ...core.integrator import NativeKernel, make_do_gradient_step_kernel

THEORY_ID = THEORY_ANISOTROPIC_SC
CHARGE_KINDS = {"vortex_number": 0}

do_gradient_step_kernel = make_do_gradient_step_kernel(THEORY_ID)
compute_energy_kernel = NativeKernel(THEORY_ID, "energy")
compute_vortex_number_kernel = NativeKernel(THEORY_ID, "charge:vortex_number")
