"""Ferromagnetic superconductor parameters (same fields, derivations and packing as
the reference's theories/ferromagnetic_superconductor/params.py: ``Params``
:20-100, ``from_params`` :190-240 -- eta2 = eta1/(-2 ha) unless given; M0, u1
from the coupled vacuum equations when hb beta - 4 eta1^2 != 0, else 1; ainf =
vortex_number/q -- and ``pack_device_params`` :266-312)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ...core.params import Params as CoreParams
from ...core.params import ResolvedParams as CoreResolvedParams
from ...core.params import pack_device_params as pack_core_device_params


@dataclass(frozen=True)
class Params(CoreParams):
    number_total_fields: int = 8
    number_magnetization_fields: int = 3
    number_higgs_fields: int = 2
    number_gauge_fields: int = 3
    q: float = 1.0
    alpha: float = -1.0
    beta: float = 1.0
    gamma: float = 1.0
    ha: float = -4.0
    hb: float = 1.0
    eta1: float = 0.0
    eta2: float | None = None
    M0: float | None = None
    u1: float | None = None
    vortex_number: float = 1.0
    ainf: float | None = None
    skyrmion_number: float = 1.0
    skyrmion_rotation: float = 0.0
    ansatz: str = "bloch"

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams(CoreResolvedParams):
    number_magnetization_fields: int
    number_higgs_fields: int
    number_gauge_fields: int
    q: float
    alpha: float
    beta: float
    gamma: float
    skyrmion_number: float
    ha: float
    hb: float
    eta1: float
    eta2: float
    M0: float
    u1: float
    vortex_number: float
    ainf: float
    skyrmion_rotation: float
    ansatz_bloch: bool
    ansatz_neel: bool
    ansatz_anti: bool
    ansatz_uniform: bool

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        eta2 = p.eta2 if p.eta2 is not None else p.eta1 / (-2.0 * p.ha)
        if p.M0 is None or p.u1 is None:
            den = p.hb * p.beta - 4.0 * p.eta1 * p.eta1
            if den != 0.0:
                M0 = math.sqrt((2.0 * p.ha * p.eta1 - p.alpha * p.hb) / den)
                u1 = math.sqrt((2.0 * p.alpha * p.eta1 - p.ha * p.beta) / den)
            else:
                M0, u1 = 1.0, 1.0
        else:
            M0, u1 = float(p.M0), float(p.u1)
        ainf = p.ainf if p.ainf is not None else p.vortex_number / p.q
        ans = (p.ansatz or "bloch").lower()
        return ResolvedParams(**CoreResolvedParams.core_fields(p),
                              number_magnetization_fields=p.number_magnetization_fields,
                              number_higgs_fields=p.number_higgs_fields, number_gauge_fields=p.number_gauge_fields,
                              q=float(p.q), alpha=float(p.alpha), beta=float(p.beta), gamma=float(p.gamma),
                              skyrmion_number=float(p.skyrmion_number), ha=float(p.ha), hb=float(p.hb),
                              eta1=float(p.eta1), eta2=float(eta2), M0=float(M0), u1=float(u1),
                              vortex_number=float(p.vortex_number), ainf=float(ainf),
                              skyrmion_rotation=float(p.skyrmion_rotation), ansatz_bloch=ans == "bloch",
                              ansatz_neel=ans == "neel", ansatz_anti=ans == "anti", ansatz_uniform=ans == "uniform")


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """Core + p_i[10..12] = nmag, nhiggs, ngauge; p_f[6..23] = q, ha, hb, eta1, eta2, u1,
    vortex_number, ainf, alpha, beta, gamma, M0, skyrmion_number, skyrmion_rotation and the
    four ansatz flags (bloch, neel, anti, uniform)."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.array([rp.number_magnetization_fields, rp.number_higgs_fields, rp.number_gauge_fields], dtype=np.int32)
    p_f = np.array([rp.q, rp.ha, rp.hb, rp.eta1, rp.eta2, rp.u1, rp.vortex_number, rp.ainf, rp.alpha, rp.beta,
                    rp.gamma, rp.M0, rp.skyrmion_number, rp.skyrmion_rotation, float(rp.ansatz_bloch),
                    float(rp.ansatz_neel), float(rp.ansatz_anti), float(rp.ansatz_uniform)], dtype=np.float64)
    return np.concatenate((p_i_core, p_i)), np.concatenate((p_f_core, p_f))


def describe() -> None:
    print("Ferromagnetic superconductor: magnetisation m (|m| = M0), gauge field A1..A3, order parameter psi.")
    print("ha, hb: quadratic/quartic |psi| coefficients; alpha, beta: quadratic/quartic |m| coefficients;")
    print("gamma: magnetic stiffness; eta1, eta2: couplings of |m|^2 and |grad m|^2 to |psi|^2.")
    print("eta2 defaults to eta1/(-2 ha); M0, u1 default to the coupled vacuum values; ainf to vortex_number/q.")
