"""Theory-independent parameters: ``Params`` -> ``ResolvedParams`` -> packed arrays.

Same fields, defaults, derivations and packed layout as the reference's
core/params.py (``Params`` :15-113, ``ResolvedParams.from_params`` :182-207,
``default_params`` :209-227, ``pack_device_params`` :229-249), because the
packed ``p_i``/``p_f`` arrays are the plugin ABI every theory extends and
that ``libsoliton.so`` reads.
"""
from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np


@dataclass(frozen=True)
class Params:
    """User-facing lattice and flow parameters shared by all theories."""

    xlen: int = 256
    ylen: int = 256
    halo: int = 2
    number_coordinates: int = 2
    number_total_fields: int = 0
    xsize: float = 80.0
    ysize: float = 80.0
    lsx: float | None = None
    lsy: float | None = None
    courant: float = 0.5
    time_step: float | None = None
    killkinen: bool = True
    newtonflow: bool = True
    unit_magnetization: bool = False

    def with_(self, **kwargs) -> "Params":
        """Copy with fields overridden (dataclasses.replace)."""
        return replace(self, **kwargs)

    def resolved(self) -> "ResolvedParams":
        return ResolvedParams.from_params(self)


@dataclass(frozen=True)
class ResolvedParams:
    """Solver-ready parameters: spacings, cell area, time step, integer flags."""

    xlen: int
    ylen: int
    halo: int
    number_coordinates: int
    number_total_fields: int
    dim_grid: int
    dim_fields: int
    killkinen: int
    newtonflow: int
    unit_magnetization: int
    xsize: float
    ysize: float
    lsx: float
    lsy: float
    grid_volume: float
    time_step: float

    @staticmethod
    def core_fields(p: Params) -> dict:
        """The derived core values (core/params.py:201-207).

        lsx = xsize/(xlen-1) unless given, likewise lsy; grid_volume = lsx*lsy;
        time_step = courant*lsx unless given (lsy plays no part).
        """
        lsx = p.lsx if p.lsx is not None else p.xsize / (p.xlen - 1)
        lsy = p.lsy if p.lsy is not None else p.ysize / (p.ylen - 1)
        time_step = p.time_step if p.time_step is not None else (p.courant * lsx)
        dim_grid = p.xlen * p.ylen
        return dict(
            xlen=p.xlen, ylen=p.ylen, halo=p.halo, number_coordinates=p.number_coordinates,
            number_total_fields=int(p.number_total_fields), dim_grid=dim_grid,
            dim_fields=int(p.number_total_fields) * dim_grid,
            killkinen=1 if p.killkinen else 0, newtonflow=1 if p.newtonflow else 0,
            unit_magnetization=1 if p.unit_magnetization else 0,
            xsize=float(p.xsize), ysize=float(p.ysize), lsx=float(lsx), lsy=float(lsy),
            grid_volume=float(lsx * lsy), time_step=float(time_step))

    @staticmethod
    def from_params(p: Params) -> "ResolvedParams":
        return ResolvedParams(**ResolvedParams.core_fields(p))


def default_params(**overrides) -> Params:
    return Params().with_(**overrides)


def pack_device_params(rp: ResolvedParams):
    """``p_i:int32[10]``, ``p_f:float64[6]`` in the reference order (:247-248).

    p_i = [xlen, ylen, halo, ncoord, nf, dim_grid, dim_fields, killkinen,
    newtonflow, unit_magnetization]; p_f = [xsize, ysize, lsx, lsy,
    grid_volume, time_step].
    """
    p_i = np.array([rp.xlen, rp.ylen, rp.halo, rp.number_coordinates, rp.number_total_fields, rp.dim_grid,
                    rp.dim_fields, rp.killkinen, rp.newtonflow, rp.unit_magnetization], dtype=np.int32)
    p_f = np.array([rp.xsize, rp.ysize, rp.lsx, rp.lsy, rp.grid_volume, rp.time_step], dtype=np.float64)
    return p_i, p_f
