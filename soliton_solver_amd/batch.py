"""Many independent relaxations of one theory in one set of launches.

``BatchSimulation(theory, [params_0, params_1, ...])`` is the batched
counterpart of a list of reference ``Simulation`` objects (one per lattice,
core/simulation.py:43-409): every instance keeps its own lattice shape,
spacing, time step, couplings and arrest state, but each RK4 pass is a single
kernel launch over the tiles of all active instances, the K-step loop is a
replayed hipGraph, and an instance retires on the device as soon as its max
|stage-4 energy gradient| drops below the tolerance.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_PREC = {"fp64": _lib.PREC_FP64, "fp32": _lib.PREC_FP32}


class BatchSimulation:
    def __init__(self, theory, params_list, precision: str = "fp32"):
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        if not params_list:
            raise ValueError("params_list is empty")
        self.theory = theory
        self.precision = precision
        self.params = list(params_list)
        self.rps = [p.resolved() for p in self.params]
        packed = [theory.params.pack_device_params(rp) for rp in self.rps]
        n_i = {len(pi) for pi, _ in packed}
        n_f = {len(pf) for _, pf in packed}
        if len(n_i) != 1 or len(n_f) != 1:
            raise ValueError("all instances must pack to the same p_i / p_f lengths")
        self.p_i = np.ascontiguousarray(np.stack([pi for pi, _ in packed]), dtype=np.int32)
        self.p_f = np.ascontiguousarray(np.stack([pf for _, pf in packed]), dtype=np.float64)
        self.ninst = len(self.params)
        b = C.c_void_p()
        _lib.check(_lib.lib().ss_batch_create(int(theory.kernels.THEORY_ID), self.ninst, self.p_i.ravel(),
                                              self.p_i.shape[1], self.p_f.ravel(), self.p_f.shape[1],
                                              _PREC[precision], C.byref(b)))
        self._b = b

    def close(self):
        if self._b:
            _lib.lib().ss_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        _lib.check(rc, batch=self._b)

    def field_shape(self, i):
        rp = self.rps[i]
        return (int(rp.number_total_fields), int(rp.xlen), int(rp.ylen))

    def grid(self, i):
        rp = self.rps[i]
        g = np.empty((2, int(rp.xlen), int(rp.ylen)))
        g[0] = rp.lsx * np.arange(int(rp.xlen))[:, None]
        g[1] = rp.lsy * np.arange(int(rp.ylen))[None, :]
        return g

    def set_field(self, i, F, V=None):
        F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).ravel())
        self._check(_lib.lib().ss_batch_upload(self._b, int(i), _lib.BUF_FIELD, F, F.size))
        V = np.zeros_like(F) if V is None else np.ascontiguousarray(np.asarray(V, dtype=np.float64).ravel())
        self._check(_lib.lib().ss_batch_upload(self._b, int(i), _lib.BUF_VELOCITY, V, V.size))

    def get_field(self, i):
        out = np.empty(int(np.prod(self.field_shape(i))))
        self._check(_lib.lib().ss_batch_download(self._b, int(i), _lib.BUF_FIELD, out, out.size))
        return out.reshape(self.field_shape(i))

    def initialize(self, configs):
        """One theory ``initial_config`` dict per instance (host-side generators);
        ``{"mode": "array", "field": F, "velocity": V}`` injects arrays, as for
        ``Simulation.initialize``."""
        for i, cfg in enumerate(configs):
            shape = self.field_shape(i)
            F = np.zeros(shape)
            V = np.zeros(shape)
            if str(cfg.get("mode", "initial")).lower() == "array":
                F[...] = np.asarray(cfg["field"], dtype=np.float64).reshape(shape)
                if cfg.get("velocity") is not None:
                    V[...] = np.asarray(cfg["velocity"], dtype=np.float64).reshape(shape)
            else:
                self.theory.initial_config.initialize(Field=F, Velocity=V, grid=self.grid(i), p_i_h=self.p_i[i],
                                                      p_f_h=self.p_f[i], config=cfg)
            self.set_field(i, F, V)
        self.begin()

    def begin(self):
        """Start (or restart) a relaxation: every instance active, its energy the
        first prev_energy, step counters at 0."""
        self._check(_lib.lib().ss_batch_begin(self._b))

    def set_precision(self, precision: str) -> None:
        """Convert every instance's field and velocity to ``precision`` in place
        ("fp32" <-> "fp64"); the arrest state is kept."""
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        self._check(_lib.lib().ss_batch_set_precision(self._b, _PREC[precision]))
        self.precision = precision

    def relax(self, tol: float, max_steps: int, check_every: int = 500, reduced_tol: float = 1e-3,
              reduced_steps: int | None = None, on_chunk=None) -> dict:
        """Relax every instance to max |EG| < ``tol``: an FP32 phase (when the
        batch is FP32) to ``reduced_tol`` -- the FP32 floor of max |EG|, set by
        the field's rounding -- for at most ``reduced_steps`` steps, then an
        FP64 retire phase (fields converted in place, every instance
        reactivated) until each reaches ``tol``; at most ``max_steps`` steps in
        all.  Returns per-instance arrays: steps of each phase, the total step
        at which the instance converged (-1: not within the budget, -2: failed
        with a non-finite error) and its last max |EG|."""
        n = self.ninst
        steps32 = np.zeros(n, dtype=np.int64)
        done = 0
        if self.precision == "fp32" and tol < reduced_tol:
            cap = max_steps if reduced_steps is None else min(reduced_steps, max_steps)
            self.begin()
            active = n
            while done < cap and active > 0:
                k = min(10 * check_every, cap - done)
                active = self.run(k, reduced_tol, check_every=check_every)
                done += k
                if on_chunk:
                    on_chunk("fp32", done, active, self.state())
            st = self.state()
            steps32 = st["steps"].astype(np.int64)
            self.set_precision("fp64")
            done = int(steps32.max()) if n else 0
        self.begin()
        active = n
        done64 = 0
        while done + done64 < max_steps and active > 0:
            k = min(10 * check_every, max_steps - done - done64)
            active = self.run(k, tol, check_every=check_every)
            done64 += k
            if on_chunk:
                on_chunk(self.precision, done + done64, active, self.state())
        st = self.state()
        conv = np.where(st["converged_step"] >= 0, st["converged_step"] + steps32, st["converged_step"])
        return {"steps_reduced": steps32, "steps_fp64": st["steps"].astype(np.int64), "converged_step": conv,
                "err_absmax": st["err_absmax"], "energy": st["energy"], "converged": int((conv >= 0).sum()),
                "failed": int((conv == -2).sum())}

    def autotune_tiles(self, reps: int = 10) -> dict:
        """Time the four RK4 passes in each shape variant (0-2 LDS tiles of the
        default rows, halved, quartered; 3-8 streaming strips, as for a single
        lattice), and with passes 2+3 as one sweep where the variant has it, on
        the live buffers and keep the fastest; the relaxation state is unchanged."""
        chosen = C.c_int()
        ms = np.zeros(27)
        self._check(_lib.lib().ss_batch_autotune_tiles(self._b, int(reps), C.byref(chosen), ms))
        info = self.tile_variant()
        info["ms"] = ms[:9]
        info["ms_fused23"] = ms[9:18]
        info["ms_pairs"] = ms[18:]
        return info

    def set_tile_variant(self, variant: int) -> None:
        self._check(_lib.lib().ss_batch_set_tile_variant(self._b, int(variant)))

    def set_fused_stages(self, fuse23: bool, pairs: bool = False) -> None:
        """Run passes 2 and 3 of every instance as one sweep, or (``pairs``,
        alone) a step as the paired sweeps 1+2 and 3+4 (fails where the current
        shape variant has no such sweep)."""
        flags = (_lib.FUSE_23 if fuse23 else 0) | (_lib.FUSE_PAIRS if pairs else 0)
        self._check(_lib.lib().ss_batch_set_fused_stages(self._b, flags))

    def tile_variant(self) -> dict:
        v, rows, f = C.c_int(), C.c_int(), C.c_int()
        self._check(_lib.lib().ss_batch_tile_variant(self._b, C.byref(v), C.byref(rows)))
        self._check(_lib.lib().ss_batch_get_fused_stages(self._b, C.byref(f)))
        return {"variant": v.value, "rows": rows.value, "fused23": bool(f.value & _lib.FUSE_23),
                "pairs": bool(f.value & _lib.FUSE_PAIRS)}

    def use_graphs(self, steps_per_graph: int):
        self._check(_lib.lib().ss_batch_use_graphs(self._b, int(steps_per_graph)))

    def run(self, nsteps: int, tol: float, check_every: int = 100) -> int:
        """Advance up to ``nsteps``; returns how many instances are still active.
        An instance whose max |EG| turns NaN or infinite retires as failed
        (``state()["failed"]``), never as converged."""
        n = C.c_int()
        self._check(_lib.lib().ss_batch_step(self._b, int(nsteps), float(tol), int(check_every), C.byref(n)))
        return n.value

    def state(self):
        E = np.empty(self.ninst)
        err = np.empty(self.ninst)
        steps = np.empty(self.ninst, dtype=np.int32)
        conv = np.empty(self.ninst, dtype=np.int32)
        self._check(_lib.lib().ss_batch_state(self._b, E, err, steps, conv))
        # converged_step: the step an instance converged at, -1 still running,
        # -2 retired with a non-finite max |EG|
        return {"energy": E, "err_absmax": err, "steps": steps, "converged_step": conv, "failed": conv == -2}

    @property
    def num_tiles(self) -> int:
        return int(_lib.lib().ss_batch_num_tiles(self._b))

    def memory_bytes(self) -> int:
        n = C.c_size_t()
        self._check(_lib.lib().ss_batch_memory_bytes(self._b, C.byref(n)))
        return int(n.value)
