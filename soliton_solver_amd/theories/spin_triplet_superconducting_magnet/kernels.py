"""Native kernel handles of the spin-triplet superconducting ferromagnet (csrc/theory_st.h,
restating the reference's theories/spin_triplet_superconducting_magnet/kernels.py)."""
from ..._lib import THEORY_SPIN_TRIPLET
from ...core.integrator import NativeKernel, make_do_gradient_step_kernel, make_do_rk4_kernel

THEORY_ID = THEORY_SPIN_TRIPLET
# vortex densities: which = 2 -> B3 = d_x A2 - d_y A1, which = 0 -> d_y A3, which = 1 -> -d_x A3
CHARGE_KINDS = {"skyrmion_number": 0, "vortex_number": 1, "vortex_number_0": 2, "vortex_number_1": 3}

do_gradient_step_kernel = make_do_gradient_step_kernel(THEORY_ID)
do_rk4_kernel = make_do_rk4_kernel(THEORY_ID)   # |m| = M0 constraint + velocity projection
compute_energy_kernel = NativeKernel(THEORY_ID, "energy")
compute_skyrmion_number_kernel = NativeKernel(THEORY_ID, "charge:skyrmion_number")
compute_vortex_number_kernel = NativeKernel(THEORY_ID, "charge:vortex_number")
