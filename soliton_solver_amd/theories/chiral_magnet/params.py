"""Chiral magnet parameters (same fields, derivations and packing as the
reference's theories/chiral_magnet/params.py: ``Params`` :21-116,
``ResolvedParams.from_params`` :199-311, ``pack_device_params`` :333-378)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable

import numpy as np

from ...core.params import Params as CoreParams
from ...core.params import ResolvedParams as CoreResolvedParams
from ...core.params import pack_device_params as pack_core_device_params

_DMI_KEYS = {
    "dresselhaus": "dresselhaus", "bulk": "dresselhaus",
    "rashba": "rashba", "interfacial": "rashba",
    "heusler": "heusler", "b20": "heusler",
    "hybrid": "hybrid",
}


@dataclass(frozen=True)
class Params(CoreParams):
    number_total_fields: int = 4
    number_magnetization_fields: int = 3
    J: float = 40e-12
    K: float = 0.8e+6
    D: float = 4e-3
    M: float = 580e+3
    B: float = 0e-3
    mu0: float = 1.25663706127e-6
    coup_K: float | None = None
    coup_h: float | None = None
    coup_mu: float | None = None
    dmi_dresselhaus: bool = True
    dmi_rashba: bool = False
    dmi_heusler: bool = False
    dmi_hybrid: bool = False
    demag: bool = False
    dmi_term: str | Iterable[str] | None = None
    skyrmion_number: float = 1.0
    skyrmion_rotation: float = 0.0
    ansatz: str = "bloch"

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams(CoreResolvedParams):
    number_magnetization_fields: int
    J: float
    K: float
    D: float
    M: float
    B: float
    mu0: float
    coup_K: float
    coup_h: float
    coup_mu: float
    skyrmion_number: float
    skyrmion_rotation: float
    ansatz_bloch: bool
    ansatz_neel: bool
    ansatz_anti: bool
    ansatz_uniform: bool
    dmi_dresselhaus: bool
    dmi_rashba: bool
    dmi_heusler: bool
    dmi_hybrid: bool
    demag: bool

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        core = CoreResolvedParams.core_fields(p)
        ans = (p.ansatz or "bloch").lower()
        # Dimensionless couplings K J/D^2, M B J/D^2, mu0 M^2 J/D^2; all 1 when D == 0.
        if p.coup_K is None or p.coup_h is None or p.coup_mu is None:
            d2 = p.D * p.D
            if d2 != 0.0:
                coup_K, coup_h, coup_mu = p.K * p.J / d2, p.M * p.B * p.J / d2, p.mu0 * p.M * p.M * p.J / d2
            else:
                coup_K = coup_h = coup_mu = 1.0
        else:
            coup_K, coup_h, coup_mu = float(p.coup_K), float(p.coup_h), float(p.coup_mu)
        flags = {"dresselhaus": bool(p.dmi_dresselhaus), "rashba": bool(p.dmi_rashba),
                 "heusler": bool(p.dmi_heusler), "hybrid": bool(p.dmi_hybrid)}
        if p.dmi_term is not None:
            flags = dict.fromkeys(flags, False)
            items = [p.dmi_term] if isinstance(p.dmi_term, str) else list(p.dmi_term)
            for item in items:
                key = str(item or "").strip().lower().replace("_", " ").replace("-", " ")
                if key not in _DMI_KEYS:
                    raise ValueError(f"Unknown DMI '{item}'. Valid: Dresselhaus (bulk), Rashba (interfacial), "
                                     "Heusler (B20), Hybrid")
                flags[_DMI_KEYS[key]] = True
            if not any(flags.values()):
                raise ValueError("p.dmi_term was provided but no valid DMIs were selected.")
        return ResolvedParams(
            **core,
            number_magnetization_fields=int(p.number_magnetization_fields),
            J=float(p.J), K=float(p.K), D=float(p.D), M=float(p.M), B=float(p.B), mu0=float(p.mu0),
            coup_K=float(coup_K), coup_h=float(coup_h), coup_mu=float(coup_mu),
            skyrmion_number=float(p.skyrmion_number), skyrmion_rotation=float(p.skyrmion_rotation),
            ansatz_bloch=ans == "bloch", ansatz_neel=ans == "neel", ansatz_anti=ans == "anti",
            ansatz_uniform=ans == "uniform",
            dmi_dresselhaus=flags["dresselhaus"], dmi_rashba=flags["rashba"],
            dmi_heusler=flags["heusler"], dmi_hybrid=flags["hybrid"], demag=bool(p.demag))


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """Core arrays + p_i[10..15] = nmag, dresselhaus, rashba, heusler, hybrid, demag;
    p_f[6..14] = coup_K, coup_h, coup_mu, skN, rotation, bloch, neel, anti, uniform."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.array([rp.number_magnetization_fields, int(rp.dmi_dresselhaus), int(rp.dmi_rashba),
                    int(rp.dmi_heusler), int(rp.dmi_hybrid), int(rp.demag)], dtype=np.int32)
    p_f = np.array([rp.coup_K, rp.coup_h, rp.coup_mu, rp.skyrmion_number, rp.skyrmion_rotation,
                    float(rp.ansatz_bloch), float(rp.ansatz_neel), float(rp.ansatz_anti),
                    float(rp.ansatz_uniform)], dtype=np.float64)
    return np.concatenate((p_i_core, p_i)), np.concatenate((p_f_core, p_f))


def describe() -> None:
    print("Chiral magnet: fields m1, m2, m3 (unit magnetisation) and the demagnetising potential psi.")
    print("Parameters J, K, D, M, B, mu0 give coup_K = K J/D^2, coup_h = M B J/D^2, coup_mu = mu0 M^2 J/D^2.")
    print("dmi_term selects Dresselhaus (bulk), Rashba (interfacial), Heusler (B20) or Hybrid; demag toggles psi.")
