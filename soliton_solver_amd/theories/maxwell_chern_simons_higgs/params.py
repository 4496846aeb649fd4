"""Maxwell-Chern-Simons-Higgs parameters (same fields, derivations and packing as the
reference's theories/maxwell_chern_simons_higgs/params.py: ``Params`` :20-82,
``from_params`` :128-151 with ainf = vortex_number/q unless given,
``pack_device_params`` :173-207)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ...core.params import Params as CoreParams
from ...core.params import ResolvedParams as CoreResolvedParams
from ...core.params import pack_device_params as pack_core_device_params


@dataclass(frozen=True)
class Params(CoreParams):
    number_total_fields: int = 5
    number_higgs_fields: int = 2
    number_gauge_fields: int = 2
    q: float = 1.0
    Lambda: float = 1.0
    kappa: float = 0.0
    u1: float = 1.0
    vortex_number: float = 1.0
    ainf: float | None = None

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams(CoreResolvedParams):
    number_higgs_fields: int
    number_gauge_fields: int
    q: float
    Lambda: float
    kappa: float
    u1: float
    vortex_number: float
    ainf: float

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        ainf = p.ainf if p.ainf is not None else (p.vortex_number / p.q)
        return ResolvedParams(**CoreResolvedParams.core_fields(p), number_higgs_fields=p.number_higgs_fields,
                              number_gauge_fields=p.number_gauge_fields, q=float(p.q), Lambda=float(p.Lambda),
                              kappa=float(p.kappa), u1=float(p.u1), vortex_number=float(p.vortex_number),
                              ainf=float(ainf))


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """Core + p_i[10..11] = nhiggs, ngauge; p_f[6..11] = q, Lambda, kappa, u1, vortex_number, ainf."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.array([rp.number_higgs_fields, rp.number_gauge_fields], dtype=np.int32)
    p_f = np.array([rp.q, rp.Lambda, rp.kappa, rp.u1, rp.vortex_number, rp.ainf], dtype=np.float64)
    return np.concatenate((p_i_core, p_i)), np.concatenate((p_f_core, p_f))


def describe() -> None:
    print("Maxwell-Chern-Simons-Higgs: gauge field A1, A2, scalar potential A0, order parameter psi; kappa is the")
    print("Chern-Simons coupling.")
