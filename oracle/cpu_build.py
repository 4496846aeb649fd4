"""Build and load the oracle's C++/OpenMP restatement -- TEST INFRASTRUCTURE ONLY.

``oracle/cpu_ref.cpp`` is compiled with g++ -O3 -fopenmp into
``oracle/_cpu/libcpuref.so`` (git-ignored, travels with the tree).  The
reference itself ships no C/C++ sources (all of it is Python + Numba), so
there is no ``oracle/_ref`` build of reference code.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_ref.cpp"
OUT = HERE / "_cpu" / "libcpuref.so"


def build() -> Path:
    OUT.parent.mkdir(exist_ok=True)
    march = os.environ.get("SS_CPU_MARCH", "-march=x86-64-v3")
    cmd = ["g++", "-O3", march, "-fopenmp", "-fPIC", "-shared", "-std=c++17", str(SRC), "-o", str(OUT)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed:\n{r.stderr}")
    return OUT


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not OUT.exists():
            build()
        L = C.CDLL(str(OUT))
        i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        f64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.cref_chiral_relax.restype = C.c_int
        L.cref_chiral_relax.argtypes = [i32, C.c_int, f64, C.c_int, f64, f64, C.c_double, C.c_int, f64, f64, C.c_int]
        L.cref_chiral_relax_tol.restype = C.c_int
        L.cref_chiral_relax_tol.argtypes = [i32, C.c_int, f64, C.c_int, f64, f64, C.c_double, C.c_int, C.c_double,
                                            f64, f64, f64, C.c_int]
        L.cref_max_threads.restype = C.c_int
        _lib = L
    return _lib


def chiral_relax(p_i, p_f, F, V, prev_energy, nsteps, nthreads=0):
    """Run ``nsteps`` reference steps in place on F, V; returns (energies, errs)."""
    L = lib()
    p_i = np.ascontiguousarray(p_i, dtype=np.int32)
    p_f = np.ascontiguousarray(p_f, dtype=np.float64)
    assert F.flags.c_contiguous and V.flags.c_contiguous and F.dtype == np.float64
    Es = np.empty(max(nsteps, 1))
    errs = np.empty(max(nsteps, 1))
    rc = L.cref_chiral_relax(p_i, p_i.size, p_f, p_f.size, F.reshape(-1), V.reshape(-1), float(prev_energy),
                             int(nsteps), Es, errs, int(nthreads))
    if rc != 0:
        raise ValueError("cref_chiral_relax rejected its parameters")
    return Es[:nsteps], errs[:nsteps]


def chiral_relax_to(p_i, p_f, F, V, prev_energy, max_steps, tol, nthreads=0):
    """Steps in place on F, V until max |stage-4 EG| < tol (at most
    ``max_steps``); returns (energies, errs, errs_abs) of the steps taken."""
    L = lib()
    p_i = np.ascontiguousarray(p_i, dtype=np.int32)
    p_f = np.ascontiguousarray(p_f, dtype=np.float64)
    assert F.flags.c_contiguous and V.flags.c_contiguous and F.dtype == np.float64
    n = max(max_steps, 1)
    Es, errs, ea = np.empty(n), np.empty(n), np.empty(n)
    done = L.cref_chiral_relax_tol(p_i, p_i.size, p_f, p_f.size, F.reshape(-1), V.reshape(-1), float(prev_energy),
                                   int(max_steps), float(tol), Es, errs, ea, int(nthreads))
    if done < 0:
        raise ValueError("cref_chiral_relax_tol rejected its parameters")
    return Es[:done], errs[:done], ea[:done]


def max_threads() -> int:
    return int(lib().cref_max_threads())
