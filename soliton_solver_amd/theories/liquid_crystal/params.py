"""Liquid crystal parameters (same fields, derivations and packing as the reference's
theories/liquid_crystal/params.py: E = voltage/d, q0 = 2 pi/P, coup_eps = K eps0/e1^2
(0 when e1 = 0), coup_Pot = 1/(q0^2 K), coup_PotE = coup_Pot eps0 delta_eps E^2,
coup_Potw0 = coup_Pot w0; a ``deformation`` ("twist" = Dresselhaus,
"splay-bend" = Rashba, or a list) overrides the two DMI flags and rejects
unknown names with ValueError)."""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Iterable

import numpy as np

from ...core.params import Params as CoreParams
from ...core.params import ResolvedParams as CoreResolvedParams
from ...core.params import pack_device_params as pack_core_device_params


@dataclass(frozen=True)
class Params(CoreParams):
    number_total_fields: int = 4
    number_magnetization_fields: int = 3
    K: float = 10e-12
    P: float = 7.0e-6
    d: float = 4e-6
    e1: float = 2e-12
    e3: float = 4e-12
    w0: float = 1.0
    eps0: float = 8.854e-12
    voltage: float = 4.0
    delta_eps: float = 3.7
    E: float | None = None
    q0: float | None = None
    coup_eps: float | None = None
    coup_Pot: float | None = None
    dmi_dresselhaus: bool = True
    dmi_rashba: bool = False
    depol: bool = False
    deformation: str | Iterable[str] | None = None
    skyrmion_number: float = 1.0
    skyrmion_rotation: float = 0.0
    ansatz: str = "bloch"

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams(CoreResolvedParams):
    number_magnetization_fields: int
    K: float
    P: float
    d: float
    e1: float
    e3: float
    w0: float
    eps0: float
    voltage: float
    delta_eps: float
    E: float
    q0: float
    coup_eps: float
    coup_Pot: float
    coup_PotE: float
    coup_Potw0: float
    skyrmion_number: float
    skyrmion_rotation: float
    ansatz_bloch: bool
    ansatz_neel: bool
    ansatz_anti: bool
    ansatz_uniform: bool
    dmi_dresselhaus: bool
    dmi_rashba: bool
    depol: bool

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        E = float(p.voltage / p.d) if p.E is None else float(p.E)
        q0 = float(2.0 * math.pi / p.P) if p.q0 is None else float(p.q0)
        if p.coup_eps is None:
            coup_eps = float(p.K * p.eps0 / (p.e1 * p.e1)) if p.e1 != 0.0 else 0.0
        else:
            coup_eps = float(p.coup_eps)
        if p.coup_Pot is None:
            coup_Pot = float(1.0 / (q0 * q0 * p.K)) if (q0 != 0.0 and p.K != 0.0) else 0.0
        else:
            coup_Pot = float(p.coup_Pot)
        dresselhaus, rashba = bool(p.dmi_dresselhaus), bool(p.dmi_rashba)
        if p.deformation is not None:
            dresselhaus = rashba = False
            items = [p.deformation] if isinstance(p.deformation, str) else list(p.deformation)
            for item in items:
                key = str(item or "").strip().lower().replace("_", " ").replace("-", " ")
                if key in ("dresselhaus", "twist"):
                    dresselhaus = True
                elif key in ("rashba", "splay bend", "splaybend"):
                    rashba = True
                else:
                    raise ValueError(f"Unknown deformation '{item}'. Valid: twist, splay-bend")
            if not (dresselhaus or rashba):
                raise ValueError("p.deformation was provided but no valid DMIs were selected.")
        ans = (p.ansatz or "bloch").lower()
        return ResolvedParams(**CoreResolvedParams.core_fields(p),
                              number_magnetization_fields=int(p.number_magnetization_fields), K=float(p.K),
                              P=float(p.P), d=float(p.d), e1=float(p.e1), e3=float(p.e3), w0=float(p.w0),
                              eps0=float(p.eps0), voltage=float(p.voltage), delta_eps=float(p.delta_eps), E=E, q0=q0,
                              coup_eps=coup_eps, coup_Pot=coup_Pot,
                              coup_PotE=float(coup_Pot * p.eps0 * p.delta_eps * E * E),
                              coup_Potw0=float(coup_Pot * p.w0), skyrmion_number=float(p.skyrmion_number),
                              skyrmion_rotation=float(p.skyrmion_rotation), ansatz_bloch=ans == "bloch",
                              ansatz_neel=ans == "neel", ansatz_anti=ans == "anti", ansatz_uniform=ans == "uniform",
                              dmi_dresselhaus=dresselhaus, dmi_rashba=rashba, depol=bool(p.depol))


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """Core + p_i[10..13] = nmag, dresselhaus, rashba, depol; p_f[6..16] = coup_PotE,
    coup_Potw0, coup_eps, e1, e3, skyrmion_number, skyrmion_rotation and the four ansatz flags."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.array([rp.number_magnetization_fields, int(rp.dmi_dresselhaus), int(rp.dmi_rashba), int(rp.depol)],
                   dtype=np.int32)
    p_f = np.array([rp.coup_PotE, rp.coup_Potw0, rp.coup_eps, rp.e1, rp.e3, rp.skyrmion_number, rp.skyrmion_rotation,
                    float(rp.ansatz_bloch), float(rp.ansatz_neel), float(rp.ansatz_anti), float(rp.ansatz_uniform)],
                   dtype=np.float64)
    return np.concatenate((p_i_core, p_i)), np.concatenate((p_f_core, p_f))


def describe() -> None:
    print("Chiral liquid crystal: director n (|n| = 1) and the depolarisation potential Phi.")
    print("K: elastic constant, P: pitch, d: cell thickness, e1, e3: flexoelectric coefficients,")
    print("voltage, delta_eps: applied field and dielectric anisotropy, w0: surface anchoring.")
    print("deformation: 'twist' (Dresselhaus) and/or 'splay-bend' (Rashba); depol switches on Phi.")
