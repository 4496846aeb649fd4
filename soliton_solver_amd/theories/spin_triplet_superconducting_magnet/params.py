"""Spin-triplet superconducting ferromagnet parameters (same fields, derivations and
packing as the reference's theories/spin_triplet_superconducting_magnet/params.py:
M0 = sqrt(-alpha/beta), u1 = u2 = sqrt(-(ha + 2 hc)/(hb1 + 2 hb2)) unless given
(1 when hb1 + 2 hb2 = 0); vortex_number = (u1^2 N1 + u2^2 N2)/(u1^2 + u2^2);
ainf = vortex_number/q)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ...core.params import Params as CoreParams
from ...core.params import ResolvedParams as CoreResolvedParams
from ...core.params import pack_device_params as pack_core_device_params


@dataclass(frozen=True)
class Params(CoreParams):
    number_total_fields: int = 10
    number_magnetization_fields: int = 3
    number_higgs_fields: int = 4
    number_gauge_fields: int = 3
    q: float = 1.0
    alpha: float = -1.0
    beta: float = 1.0
    gamma: float = 1.0
    ha: float = -4.0
    hb1: float = 3.0
    hb2: float = 0.5
    hc: float = -0.5
    M0: float | None = None
    u1: float | None = None
    u2: float | None = None
    vortex_number: float = 1.0
    vortex1_number: float = 1.0
    vortex2_number: float = 1.0
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
    hb1: float
    hb2: float
    hc: float
    M0: float
    u1: float
    u2: float
    vortex_number: float
    vortex1_number: float
    vortex2_number: float
    ainf: float
    skyrmion_rotation: float
    ansatz_bloch: bool
    ansatz_neel: bool
    ansatz_anti: bool
    ansatz_uniform: bool

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        if p.M0 is None or p.u1 is None or p.u2 is None:
            den = p.hb1 + 2.0 * p.hb2
            if den != 0.0:
                M0 = math.sqrt(-p.alpha / p.beta)
                u1 = u2 = math.sqrt(-(p.ha + 2.0 * p.hc) / den)
            else:
                M0 = u1 = u2 = 1.0
        else:
            M0, u1, u2 = float(p.M0), float(p.u1), float(p.u2)
        vn = (u1 * u1 * p.vortex1_number + u2 * u2 * p.vortex2_number) / (u1 * u1 + u2 * u2)
        ainf = p.ainf if p.ainf is not None else vn / p.q
        ans = (p.ansatz or "bloch").lower()
        return ResolvedParams(**CoreResolvedParams.core_fields(p),
                              number_magnetization_fields=p.number_magnetization_fields,
                              number_higgs_fields=p.number_higgs_fields, number_gauge_fields=p.number_gauge_fields,
                              q=float(p.q), alpha=float(p.alpha), beta=float(p.beta), gamma=float(p.gamma),
                              skyrmion_number=float(p.skyrmion_number), ha=float(p.ha), hb1=float(p.hb1),
                              hb2=float(p.hb2), hc=float(p.hc), M0=float(M0), u1=float(u1), u2=float(u2),
                              vortex_number=float(vn), vortex1_number=float(p.vortex1_number),
                              vortex2_number=float(p.vortex2_number), ainf=float(ainf),
                              skyrmion_rotation=float(p.skyrmion_rotation), ansatz_bloch=ans == "bloch",
                              ansatz_neel=ans == "neel", ansatz_anti=ans == "anti", ansatz_uniform=ans == "uniform")


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """Core + p_i[10..12] = nmag, nhiggs, ngauge; p_f[6..26] = q, ha, hb1, hb2, hc, u1, u2,
    vortex1_number, vortex2_number, vortex_number, ainf, alpha, beta, gamma, M0,
    skyrmion_number, skyrmion_rotation and the four ansatz flags."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.array([rp.number_magnetization_fields, rp.number_higgs_fields, rp.number_gauge_fields], dtype=np.int32)
    p_f = np.array([rp.q, rp.ha, rp.hb1, rp.hb2, rp.hc, rp.u1, rp.u2, rp.vortex1_number, rp.vortex2_number,
                    rp.vortex_number, rp.ainf, rp.alpha, rp.beta, rp.gamma, rp.M0, rp.skyrmion_number,
                    rp.skyrmion_rotation, float(rp.ansatz_bloch), float(rp.ansatz_neel), float(rp.ansatz_anti),
                    float(rp.ansatz_uniform)], dtype=np.float64)
    return np.concatenate((p_i_core, p_i)), np.concatenate((p_f_core, p_f))


def describe() -> None:
    print("Spin-triplet superconducting ferromagnet: magnetisation m (|m| = M0), gauge field A1..A3 and two")
    print("complex order parameters psi1, psi2 (equal-spin triplet pairing).")
    print("ha, hb1, hb2, hc: order-parameter potential; alpha, beta, gamma: magnetic sector.")
