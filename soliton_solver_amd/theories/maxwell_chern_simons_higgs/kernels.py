"""Native kernel handles of the Maxwell-Chern-Simons-Higgs model (csrc/theory_mcsh.h,
restating the reference's theories/maxwell_chern_simons_higgs/kernels.py)."""
from ..._lib import THEORY_MAXWELL_CHERN_SIMONS_HIGGS
from ...core.integrator import NativeKernel, make_do_gradient_step_kernel

THEORY_ID = THEORY_MAXWELL_CHERN_SIMONS_HIGGS
CHARGE_KINDS = {"vortex_number": 0, "electric_charge": 1, "noether_charge": 2}

do_gradient_step_kernel = make_do_gradient_step_kernel(THEORY_ID)
compute_energy_kernel = NativeKernel(THEORY_ID, "energy")
compute_vortex_number_kernel = NativeKernel(THEORY_ID, "charge:vortex_number")
compute_electric_charge_kernel = NativeKernel(THEORY_ID, "charge:electric_charge")
compute_noether_charge_kernel = NativeKernel(THEORY_ID, "charge:noether_charge")
