"""Native kernel handles of the chiral magnet (kernels compiled into libsoliton.so
from csrc/theory_chiral.h, restating the reference's theories/chiral_magnet/kernels.py)."""
from ..._lib import THEORY_CHIRAL_MAGNET
from ...core.integrator import NativeKernel, make_do_gradient_step_kernel, make_do_rk4_kernel

THEORY_ID = THEORY_CHIRAL_MAGNET
CHARGE_KINDS = {"skyrmion_number": 0, "magnetic_charge": 1}

do_gradient_step_kernel = make_do_gradient_step_kernel(THEORY_ID)
do_rk4_kernel = make_do_rk4_kernel(THEORY_ID)   # unit-norm constraint + velocity projection
compute_energy_kernel = NativeKernel(THEORY_ID, "energy")
compute_skyrmion_number_kernel = NativeKernel(THEORY_ID, "charge:skyrmion_number")
compute_magnetic_charge_kernel = NativeKernel(THEORY_ID, "charge:magnetic_charge")
