"""Baby Skyrme parameters (same fields, derivations and packing as the reference's
theories/baby_skyrme/params.py: ``Params`` :24-113, ``from_params`` :192-307 with
the ``potential`` selector, ``pack_device_params`` :331-378)."""
from __future__ import annotations

from dataclasses import dataclass
from typing import Iterable

import numpy as np

from ...core.params import Params as CoreParams
from ...core.params import ResolvedParams as CoreResolvedParams
from ...core.params import pack_device_params as pack_core_device_params

_POT_ORDER = ("Standard", "Holomorphic", "EasyPlane", "Dihedral2", "Aloof", "Dihedral3", "Broken", "DoubleVacua")
_POT_KEYS = {
    "standard": "Standard",
    "holomorphic": "Holomorphic", "holomorhpic": "Holomorphic",
    "easy plane": "EasyPlane", "easyplane": "EasyPlane",
    "dihedral2": "Dihedral2", "dihedral 2": "Dihedral2",
    "aloof": "Aloof",
    "dihedral3": "Dihedral3", "dihedral 3": "Dihedral3",
    "broken": "Broken",
    "double vacua": "DoubleVacua", "doublevacua": "DoubleVacua", "2 vacua": "DoubleVacua", "two vacua": "DoubleVacua",
}


@dataclass(frozen=True)
class Params(CoreParams):
    number_total_fields: int = 3
    number_magnetization_fields: int = 3
    mpi: float = 1.0
    kappa: float = 1.0
    N: int = 1
    Potential_Standard: bool = True
    Potential_Holomorphic: bool = False
    Potential_EasyPlane: bool = False
    Potential_Dihedral2: bool = False
    Potential_Aloof: bool = False
    Potential_Dihedral3: bool = False
    Potential_Broken: bool = False
    Potential_DoubleVacua: bool = False
    potential: str | Iterable[str] | None = None
    skyrmion_number: float = 1.0
    skyrmion_rotation: float = 0.0
    ansatz: str = "bloch"

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams(CoreResolvedParams):
    number_magnetization_fields: int
    N: int
    mpi: float
    kappa: float
    skyrmion_number: float
    skyrmion_rotation: float
    ansatz_bloch: bool
    ansatz_neel: bool
    ansatz_anti: bool
    ansatz_uniform: bool
    Potential_Standard: bool
    Potential_Holomorphic: bool
    Potential_EasyPlane: bool
    Potential_Dihedral2: bool
    Potential_Aloof: bool
    Potential_Dihedral3: bool
    Potential_Broken: bool
    Potential_DoubleVacua: bool

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        ans = (p.ansatz or "bloch").lower()
        pots = {k: bool(getattr(p, "Potential_" + k)) for k in _POT_ORDER}
        if p.potential is not None:
            pots = dict.fromkeys(pots, False)
            items = [p.potential] if isinstance(p.potential, str) else list(p.potential)
            for item in items:
                key = str(item or "").strip().lower().replace("_", " ").replace("-", " ")
                if key not in _POT_KEYS:
                    raise ValueError(f"Unknown potential '{item}'. Valid: Standard, Holomorphic, Easy plane, "
                                     "Dihedral2, Aloof, Dihedral3, Broken, DoubleVacua")
                pots[_POT_KEYS[key]] = True
            if not any(pots.values()):
                raise ValueError("p.potential was provided but no valid potentials were selected.")
        return ResolvedParams(
            **CoreResolvedParams.core_fields(p),
            number_magnetization_fields=int(p.number_magnetization_fields), N=int(p.N),
            mpi=float(p.mpi), kappa=float(p.kappa), skyrmion_number=float(p.skyrmion_number),
            skyrmion_rotation=float(p.skyrmion_rotation),
            ansatz_bloch=ans == "bloch", ansatz_neel=ans == "neel", ansatz_anti=ans == "anti",
            ansatz_uniform=ans == "uniform",
            **{"Potential_" + k: v for k, v in pots.items()})


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """Core + p_i[10..19] = nmag, 8 potential flags, N; p_f[6..13] = mpi, kappa, skN,
    rotation, bloch, neel, anti, uniform."""
    p_i_core, p_f_core = pack_core_device_params(rp)
    p_i = np.array([rp.number_magnetization_fields] + [int(getattr(rp, "Potential_" + k)) for k in _POT_ORDER]
                   + [int(rp.N)], dtype=np.int32)
    p_f = np.array([rp.mpi, rp.kappa, rp.skyrmion_number, rp.skyrmion_rotation, float(rp.ansatz_bloch),
                    float(rp.ansatz_neel), float(rp.ansatz_anti), float(rp.ansatz_uniform)], dtype=np.float64)
    return np.concatenate((p_i_core, p_i)), np.concatenate((p_f_core, p_f))


def describe() -> None:
    print("Baby Skyrme model: unit 3-vector m; mpi (pion mass), kappa (Skyrme coupling), potential selector.")
