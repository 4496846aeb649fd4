"""Bose-Einstein condensate parameters (same fields, derivations and packing as the
reference's theories/bose_einstein_condensate/params.py: ``Params`` :19-89,
``from_params`` :147-176 -- beta = 4 pi N a_s sqrt(m omega/hbar) and
tf_radius = sqrt(2 sqrt(beta/pi)) unless given -- ``pack_device_params`` :198-229)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ...core.params import Params as CoreParams
from ...core.params import ResolvedParams as CoreResolvedParams
from ...core.params import pack_device_params as pack_core_device_params

_A0 = 0.529e-10


@dataclass(frozen=True)
class Params(CoreParams):
    number_total_fields: int = 2
    number_higgs_fields: int = 2
    N: float = 1000.0
    a_0: float = _A0
    a_s: float = 109.0 * _A0
    m: float = 1.45e-25
    omega: float = 674.2 / (2.0 * math.pi)
    omega_rot: float = 0.99
    hbar: float = 1.054571817e-34
    beta: float | None = None
    tf_radius: float | None = None
    vortex_number: float = 1.0

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams(CoreResolvedParams):
    number_higgs_fields: int
    N: float
    a_0: float
    a_s: float
    m: float
    omega: float
    hbar: float
    omega_rot: float
    beta: float
    tf_radius: float
    vortex_number: float = 1.0

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        beta = p.beta if p.beta is not None else (4.0 * math.pi * p.N * p.a_s * math.sqrt(p.m * p.omega / p.hbar))
        tf = p.tf_radius if p.tf_radius is not None else math.sqrt(2.0 * math.sqrt(beta / math.pi))
        return ResolvedParams(**CoreResolvedParams.core_fields(p), number_higgs_fields=p.number_higgs_fields,
                              N=float(p.N), a_0=float(p.a_0), a_s=float(p.a_s), m=float(p.m), omega=float(p.omega),
                              hbar=float(p.hbar), omega_rot=float(p.omega_rot), beta=float(beta),
                              tf_radius=float(tf), vortex_number=float(p.vortex_number))


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """Core + p_i[10] = nhiggs; p_f[6..9] = beta, tf_radius, omega_rot, vortex_number."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.array([rp.number_higgs_fields], dtype=np.int32)
    p_f = np.array([rp.beta, rp.tf_radius, rp.omega_rot, rp.vortex_number], dtype=np.float64)
    return np.concatenate((p_i_core, p_i)), np.concatenate((p_f_core, p_f))


def describe() -> None:
    print("Rotating Bose-Einstein condensate: psi = psi_r + i psi_i, harmonic trap, beta |psi|^4, rotation omega_rot;")
    print("the flow keeps sum |psi|^2 gv = 1.")
