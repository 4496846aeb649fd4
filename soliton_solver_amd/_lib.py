"""ctypes binding of ``libsoliton.so`` (declarations in include/soliton.h).

The library is built in-tree by ``python -m soliton_solver_amd.build``.  There
is no fallback: if the library is missing every entry point raises, so a run
without the HIP path fails loudly instead of silently computing elsewhere.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libsoliton.so"

THEORY_CHIRAL_MAGNET = 1
THEORY_GINZBURG_LANDAU = 2
THEORY_BABY_SKYRME = 3
THEORY_ANISOTROPIC_SC = 4
THEORY_BEC = 5
THEORY_MAXWELL_CHERN_SIMONS_HIGGS = 6
THEORY_FERROMAGNETIC_SC = 7
THEORY_SPIN_TRIPLET = 8
THEORY_LIQUID_CRYSTAL = 9

PREC_FP64 = 0
PREC_FP32 = 1
PREC_FP32C = 2
VAR_REFERENCE = 0
VAR_EXACT_ADJOINT = 1
AUX_HIGGS, AUX_HIGGS1, AUX_HIGGS2, AUX_PHASE_DIFFERENCE, AUX_SUPERCURRENT, AUX_MAGNETIC_FIELD = range(6)
FUSE_23 = 1
FUSE_41 = 2
FUSE_PAIRS = 4
FUSE_COMP12 = 8
BUF_FIELD = 0
BUF_VELOCITY = 1

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_dp = C.POINTER(C.c_double)

INIT_GROUND, INIT_INITIAL, INIT_SKYRMIONS, INIT_VORTEX = range(4)


class InitDesc(C.Structure):
    """``ss_init_desc``: what ss_init_field generates on the device."""

    _fields_ = [
        ("mode", C.c_int),
        ("ansatz", C.c_int),
        ("skyrmion_number", C.c_double),
        ("rotation", C.c_double),
        ("scale", C.c_double),
        ("vortex_number", C.c_double),
        ("extra", C.c_double * 4),
        ("nsol", C.c_int),
        ("solitons", C.POINTER(C.c_double)),
    ]


# name -> (restype, argtypes); mirrors include/soliton.h one to one.
SIGNATURES = {
    "ss_version": (C.c_int, []),
    "ss_global_error": (C.c_char_p, []),
    "ss_device_count": (C.c_int, []),
    "ss_ctx_create": (C.c_int, [C.c_int, _i32p, C.c_int, _f64p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "ss_ctx_destroy": (C.c_int, [C.c_void_p]),
    "ss_set_params": (C.c_int, [C.c_void_p, _i32p, C.c_int, _f64p, C.c_int]),
    "ss_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "ss_set_variation_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "ss_upload": (C.c_int, [C.c_void_p, C.c_int, _f64p, C.c_size_t]),
    "ss_download": (C.c_int, [C.c_void_p, C.c_int, _f64p, C.c_size_t]),
    "ss_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int, _dp, _dp, _dp]),
    "ss_step_history": (C.c_int, [C.c_void_p, _f64p, _f64p, C.c_int]),
    "ss_energy": (C.c_int, [C.c_void_p, _dp]),
    "ss_charge": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "ss_num_charges": (C.c_int, [C.c_void_p]),
    "ss_density": (C.c_int, [C.c_void_p, C.c_int, _f64p, C.c_size_t]),
    "ss_variation": (C.c_int, [C.c_void_p, C.c_int, _f64p, C.c_size_t]),
    "ss_variation_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "ss_energy_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ss_energy_dev_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ss_energy_param_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "ss_use_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "ss_prepare_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "ss_synchronize": (C.c_int, [C.c_void_p]),
    "ss_time_passes": (C.c_int, [C.c_void_p, C.c_double, C.c_int, _f64p]),
    "ss_set_tile_variants": (C.c_int, [C.c_void_p, _i32p]),
    "ss_get_tile_variants": (C.c_int, [C.c_void_p, _i32p, _i32p]),
    "ss_autotune_tiles": (C.c_int, [C.c_void_p, C.c_int, _i32p, _f64p]),
    "ss_set_fused_stages": (C.c_int, [C.c_void_p, C.c_int]),
    "ss_init_field": (C.c_int, [C.c_void_p, C.POINTER(InitDesc)]),
    "ss_add_noise": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_double]),
    "ss_aux_components": (C.c_int, [C.c_void_p, C.c_int]),
    "ss_aux_density": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _f64p, C.c_size_t]),
    "ss_get_fused_stages": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "ss_memory_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "ss_last_error": (C.c_char_p, [C.c_void_p]),
    "ss_batch_create": (C.c_int, [C.c_int, C.c_int, _i32p, C.c_int, _f64p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "ss_batch_destroy": (C.c_int, [C.c_void_p]),
    "ss_batch_last_error": (C.c_char_p, [C.c_void_p]),
    "ss_batch_num_tiles": (C.c_int, [C.c_void_p]),
    "ss_batch_upload": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _f64p, C.c_size_t]),
    "ss_batch_download": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _f64p, C.c_size_t]),
    "ss_batch_begin": (C.c_int, [C.c_void_p]),
    "ss_batch_set_precision": (C.c_int, [C.c_void_p, C.c_int]),
    "ss_batch_step": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int)]),
    "ss_batch_use_graphs": (C.c_int, [C.c_void_p, C.c_int]),
    "ss_batch_state": (C.c_int, [C.c_void_p, _f64p, _f64p, _i32p, _i32p]),
    "ss_batch_memory_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_size_t)]),
    "ss_batch_autotune_tiles": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int), _f64p]),
    "ss_batch_set_tile_variant": (C.c_int, [C.c_void_p, C.c_int]),
    "ss_batch_tile_variant": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ss_batch_set_fused_stages": (C.c_int, [C.c_void_p, C.c_int]),
    "ss_batch_get_fused_stages": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
}

_lib = None


class LibraryMissing(RuntimeError):
    """libsoliton.so is not built (or failed to load)."""


def _share_torch_hip_runtime() -> None:
    """Load torch's HIP runtime first, when torch is installed.

    torch ships its own libamdhip64 (soname libamdhip64.so.7, the library's
    own dependency).  Loaded first, it satisfies the library's dependency and
    the process has one HIP runtime; loaded after /opt/rocm's copy (which the
    library would pull in otherwise), torch starts a second runtime that finds
    no GPU.  The autograd boundary hands torch tensors to the library, so both
    must share one runtime.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load the library once; raise ``LibraryMissing`` if it is not there."""
    global _lib
    if _lib is not None:
        return _lib
    _share_torch_hip_runtime()
    if not LIB_PATH.exists():
        raise LibraryMissing(
            f"{LIB_PATH} is not built; run `python -m soliton_solver_amd.build` "
            "(the HIP path has no CPU fallback)")
    try:
        L = C.CDLL(str(LIB_PATH))
    except OSError as exc:
        raise LibraryMissing(f"cannot load {LIB_PATH}: {exc}") from exc
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def error_text(ctx=None, batch=None) -> str:
    L = lib()
    if batch:
        msg = L.ss_batch_last_error(batch)
    else:
        msg = L.ss_last_error(ctx) if ctx else L.ss_global_error()
    return (msg or b"").decode(errors="replace")


def check(rc: int, ctx=None, batch=None) -> None:
    """Map a non-zero status to the exception the reference raises for it.

    -1 (bad argument or unsupported request) -> ValueError,
    -2 (HIP runtime failure) -> RuntimeError.
    """
    if rc == 0:
        return
    msg = error_text(ctx, batch)
    if rc == -1:
        raise ValueError(msg)
    raise RuntimeError(msg)


def device_count() -> int:
    return int(lib().ss_device_count())


def out_double():
    return C.c_double(0.0)
