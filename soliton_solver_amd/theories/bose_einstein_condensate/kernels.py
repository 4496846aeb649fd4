"""Native kernel handles of the Bose-Einstein condensate (csrc/theory_bec.h, restating the
reference's theories/bose_einstein_condensate/kernels.py)."""
from ..._lib import THEORY_BEC
from ...core.integrator import NativeKernel, make_do_gradient_step_kernel

THEORY_ID = THEORY_BEC
CHARGE_KINDS = {"norm": 0}

do_gradient_step_kernel = make_do_gradient_step_kernel(THEORY_ID)
compute_energy_kernel = NativeKernel(THEORY_ID, "energy")
compute_norm_kernel = NativeKernel(THEORY_ID, "charge:norm")
do_norm_kernel = NativeKernel(THEORY_ID, "l2_normalize")
