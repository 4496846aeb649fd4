"""``Simulation``: the driver object of the reference API, over libsoliton.so.

Mirrors soliton_solver/core/simulation.py (``__init__`` :43-115,
``initialize`` :117-142, ``set_params`` :144-166, ``observables`` :168-192,
``step`` :388-409, ``compute_energy_density`` :194-209): same construction
from ``(params, theory)``, same plugin attribute checks, same return values.
The difference is where the state lives: the reference keeps a dozen FP64
Numba device arrays per field and runs ~15 launches and 2 host round trips per
step; here one native context (``ss_ctx``) owns the device state and a step is
four fused HIP passes with the arrest decision taken on the device.

Extras beyond the reference (no counterpart there): ``steps`` (many steps per
call), ``minimize`` (the headless loop the reference documents at
core/simulation.py:8,39 but does not define), ``precision`` ("fp64", the
"fp32" stencil / FP64-reduction form, or "fp32c", the same with a compensated
field update that keeps increments below an FP32 ulp), ``energy_gradient`` and array
injection ``initialize({"mode": "array", ...})``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from .params import Params

_PREC = {"fp64": _lib.PREC_FP64, "fp32": _lib.PREC_FP32, "fp32c": _lib.PREC_FP32C}
_VAR = {"reference": _lib.VAR_REFERENCE, "exact_adjoint": _lib.VAR_EXACT_ADJOINT}


class DeviceBuffer:
    """Host view of one device buffer with the Numba-style copy methods.

    The reference exposes ``sim.Field`` / ``sim.Velocity`` as Numba device
    arrays; code that calls ``sim.Field.copy_to_host()`` keeps working.
    Values are FP64 in the reference flat layout (core/utils.py:16-42).
    """

    def __init__(self, sim: "Simulation", buf: int):
        self._sim = sim
        self._buf = buf

    @property
    def size(self) -> int:
        return int(self._sim.rp.dim_fields)

    @property
    def shape(self):
        return (self.size,)

    def copy_to_host(self) -> np.ndarray:
        out = np.empty(self.size, dtype=np.float64)
        _lib.check(_lib.lib().ss_download(self._sim._ctx, self._buf, out, out.size), self._sim._ctx)
        return out

    def copy_to_device(self, host) -> None:
        a = np.ascontiguousarray(np.asarray(host, dtype=np.float64).ravel())
        if a.size != self.size:
            raise ValueError(f"expected {self.size} values, got {a.size}")
        _lib.check(_lib.lib().ss_upload(self._sim._ctx, self._buf, a, a.size), self._sim._ctx)

    def __array__(self, dtype=None, copy=None):
        a = self.copy_to_host()
        return a if dtype is None else a.astype(dtype)


def as_simulation(obj) -> "Simulation":
    """The Simulation behind ``obj``: a Simulation, or its ``Field`` / ``Velocity``
    buffer (the first argument of the reference's observables functions, e.g.
    ``compute_energy(Field, d1fd1x, en, entmp, gridsum_partial, p_i_d, p_f_d,
    p_i_h, p_f_h)`` in chiral_magnet/observables.py:15-57)."""
    if isinstance(obj, Simulation):
        return obj
    if isinstance(obj, DeviceBuffer):
        return obj._sim
    raise TypeError(f"expected a Simulation or its Field buffer, got {type(obj).__name__}")


def which_argument(args, default: int = 2) -> int:
    """``which`` of compute_vortex_number: the only extra argument in
    ``f(sim, which)``, the fifth in the reference's
    ``f(Field, d1fd1x, en, entmp, gridsum_partial, which, p_i_d, p_f_d, p_i_h)``."""
    if len(args) == 1:
        return int(args[0])
    if len(args) >= 5:
        return int(args[4])
    return default


def _check_finite(E: float, err: float, steps: int) -> None:
    """Raise when a relaxation has produced a non-finite energy or error."""
    if not (np.isfinite(E) and np.isfinite(err)):
        raise FloatingPointError(f"relaxation diverged: energy {E!r}, max|EG| {err!r} after {steps} steps")


class Simulation:
    """Arrested-Newton-flow relaxation of one theory on one GPU."""

    def __init__(self, params, theory, *, precision: str = "fp64", variation: str = "reference"):
        self.theory = theory
        self.kernels = theory.kernels
        self.initial_config = theory.initial_config
        self.observables_mod = theory.observables
        self.threads2d = getattr(self.theory, "threads2d", (8, 8))
        self.gradient_step_kernel = getattr(self.kernels, "do_gradient_step_kernel", None)
        if self.gradient_step_kernel is None:
            raise AttributeError("Theory kernels must provide do_gradient_step_kernel "
                                 "(a soliton_solver_amd.core.integrator.NativeKernel).")
        self.rk4_kernel = getattr(self.kernels, "do_rk4_kernel", None)
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        if variation not in _VAR:
            raise ValueError(f"variation must be one of {sorted(_VAR)}")
        self.precision = precision
        self.variation = variation
        self._ctx = None
        self.set_params(params)
        self.Field = DeviceBuffer(self, _lib.BUF_FIELD)
        self.Velocity = DeviceBuffer(self, _lib.BUF_VELOCITY)
        # names of the reference's scratch buffers, for code that passes them
        # to observables functions; the derivative and reduction scratch live
        # in the native context
        self.d1fd1x = self.d2fd2x = self.entmp = self.gridsum_partial = None
        self.en = None
        self.grid = None
        self.last_history = None

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if self._ctx:
            _lib.lib().ss_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        _lib.check(rc, self._ctx)

    # -- parameters ---------------------------------------------------------
    def set_params(self, p) -> None:
        """Resolve, pack and push parameters (core/simulation.py:144-166)."""
        self.p = p
        self.rp = p if not isinstance(p, Params) else p.resolved()
        p_i, p_f = self.theory.params.pack_device_params(self.rp)
        self.p_i_h = np.ascontiguousarray(p_i, dtype=np.int32)
        self.p_f_h = np.ascontiguousarray(p_f, dtype=np.float64)
        # the reference's device copies; the native context holds its own
        self.p_i_d, self.p_f_d = self.p_i_h, self.p_f_h
        L = _lib.lib()
        if self._ctx is None:
            ctx = C.c_void_p()
            rc = L.ss_ctx_create(int(self.kernels.THEORY_ID), self.p_i_h, self.p_i_h.size, self.p_f_h,
                                 self.p_f_h.size, _PREC[self.precision], _VAR[self.variation], C.byref(ctx))
            _lib.check(rc, None)
            self._ctx = ctx
        else:
            self._check(L.ss_set_params(self._ctx, self.p_i_h, self.p_i_h.size, self.p_f_h, self.p_f_h.size))

    def set_precision(self, precision: str) -> None:
        """Switch storage precision in place (e.g. FP32 relaxation -> FP64 polish).
        "fp64" -> "fp32c" splits the field exactly into FP32 hi + lo words and
        "fp32c" -> "fp64" joins them; "fp32" -> "fp32c" starts lo at zero."""
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {sorted(_PREC)}")
        self._check(_lib.lib().ss_set_precision(self._ctx, _PREC[precision]))
        self.precision = precision

    def use_graphs(self, steps_per_graph: int) -> None:
        """Replay ``steps_per_graph`` captured steps per hipGraph launch (0: off)."""
        self._check(_lib.lib().ss_use_graphs(self._ctx, int(steps_per_graph)))

    def prepare_graphs(self, nsteps: int) -> None:
        """Build now what the first ``steps(E, nsteps)`` would build inside the call
        (history buffer, flow scratch, the step graphs of both buffer parities),
        without running a step."""
        self._check(_lib.lib().ss_prepare_graphs(self._ctx, int(nsteps)))

    # -- state --------------------------------------------------------------
    @property
    def field_shape(self):
        return (int(self.rp.number_total_fields), int(self.rp.xlen), int(self.rp.ylen))

    def get_field(self) -> np.ndarray:
        return self.Field.copy_to_host().reshape(self.field_shape)

    def get_velocity(self) -> np.ndarray:
        return self.Velocity.copy_to_host().reshape(self.field_shape)

    def set_field(self, field) -> None:
        self.Field.copy_to_device(field)

    def set_velocity(self, velocity) -> None:
        self.Velocity.copy_to_device(velocity)

    def create_grid(self) -> np.ndarray:
        """``grid[0] = lsx*x``, ``grid[1] = lsy*y`` (e.g. chiral_magnet/kernels.py:17-45)."""
        X, Y = int(self.rp.xlen), int(self.rp.ylen)
        g = np.empty((2, X, Y))
        g[0] = self.rp.lsx * np.arange(X, dtype=np.float64)[:, None]
        g[1] = self.rp.lsy * np.arange(Y, dtype=np.float64)[None, :]
        return g

    def initialize(self, init_config: dict | None = None) -> None:
        """Build the initial state and upload it (core/simulation.py:117-142).

        ``{"mode": "array", "field": F, "velocity": V}`` injects host arrays
        (shape ``(nf, xlen, ylen)`` or flat); every other mode is passed to the
        theory's ``initial_config.initialize``, which fills host arrays in place.
        With ``"device": True`` the same configuration is generated on the
        device instead (no host field, no upload); the noise then comes from
        the device's counter-based generator seeded with ``"seed"``, not from
        NumPy's, so only noise-free configurations agree with the host ones.
        """
        cfg = dict(init_config or {})
        if cfg.get("device") and str(cfg.get("mode", "initial")).lower() != "array":
            self.initialize_on_device(cfg)
            return
        self.grid = self.create_grid()
        F = np.zeros(self.field_shape)
        V = np.zeros(self.field_shape)
        mode = str(cfg.get("mode", "initial")).lower()
        if mode == "array":
            F[...] = np.asarray(cfg["field"], dtype=np.float64).reshape(self.field_shape)
            if cfg.get("velocity") is not None:
                V[...] = np.asarray(cfg["velocity"], dtype=np.float64).reshape(self.field_shape)
        else:
            self.initial_config.initialize(Field=F, Velocity=V, grid=self.grid, p_i_h=self.p_i_h,
                                           p_f_h=self.p_f_h, config=cfg)
        self.set_field(F)
        self.set_velocity(V)

    def initialize_on_device(self, init_config: dict | None = None) -> None:
        """Generate the theory's initial configuration on the device (ss_init_field).

        The theory's ``initial_config.device_plan`` translates the host
        configuration into an ``ss_init_desc`` plus a list of noise passes
        ``(comp_begin, comp_end, gaussian, amplitude, renorm_ncomp, renorm_norm)``.
        """
        cfg = dict(init_config or {})
        plan = getattr(self.initial_config, "device_plan", None)
        if plan is None:
            raise NotImplementedError(f"theory {self.theory.THEORY_SPEC.name!r} has no device initial conditions")
        self.grid = self.create_grid()
        desc_kw, noise = plan(cfg, self.p_i_h, self.p_f_h)
        d = _lib.InitDesc()
        d.mode = int(desc_kw.get("mode", _lib.INIT_INITIAL))
        d.ansatz = int(desc_kw.get("ansatz", 0))
        d.skyrmion_number = float(desc_kw.get("skyrmion_number", 1.0))
        d.rotation = float(desc_kw.get("rotation", 0.0))
        d.scale = float(desc_kw.get("scale", 1.0))
        d.vortex_number = float(desc_kw.get("vortex_number", 1.0))
        for k, v in enumerate(desc_kw.get("extra", ())):
            d.extra[k] = float(v)
        sol = np.ascontiguousarray(np.asarray(desc_kw.get("solitons", ()), dtype=np.float64).reshape(-1))
        d.nsol = sol.size // 3
        d.solitons = sol.ctypes.data_as(C.POINTER(C.c_double)) if d.nsol else None
        L = _lib.lib()
        self._check(L.ss_init_field(self._ctx, C.byref(d)))
        seed = int(cfg.get("seed", 0))
        for k, (c0, c1, gauss, amp, rn, rv) in enumerate(noise):
            self._check(L.ss_add_noise(self._ctx, int(c0), int(c1), int(gauss), float(amp),
                                       (seed + 0x9E3779B97F4A7C15 * k) & 0xFFFFFFFFFFFFFFFF, int(rn), float(rv)))

    # -- the flow -----------------------------------------------------------
    def step(self, prev_energy: float) -> tuple[float, float]:
        """One arrested Newton flow step; returns ``(new_energy, err)``.

        core/simulation.py:388-409 -> core/integrator.py:217-348.
        """
        e, err, _ = self.steps(prev_energy, 1)
        return e, err

    def steps(self, prev_energy: float, nsteps: int, history: bool = False):
        """``nsteps`` steps in one native call, chained like the reference loop.

        Equivalent to ``E = prev_energy; for _ in range(n): E, err = sim.step(E)``
        (chiral_magnet/render_gl.py:663-668).  Returns ``(E, err, err_absmax)``
        of the last step; with ``history`` the per-step energies and errs are
        kept in ``self.last_history``.
        """
        e, em, ea = C.c_double(), C.c_double(), C.c_double()
        L = _lib.lib()
        self._check(L.ss_step(self._ctx, float(prev_energy), int(nsteps), C.byref(e), C.byref(em), C.byref(ea)))
        if history and nsteps > 0:
            hE = np.empty(nsteps)
            hR = np.empty(nsteps)
            self._check(L.ss_step_history(self._ctx, hE, hR, int(nsteps)))
            self.last_history = (hE, hR)
        return e.value, em.value, ea.value

    def minimize(self, tol: float = 1e-6, max_steps: int = 100000, check_every: int = 100,
                 prev_energy: float | None = None):
        """Relax until max |EG| < tol (checked every ``check_every`` steps).

        Returns ``(energy, err_absmax, steps_taken)``.  Raises
        ``FloatingPointError`` as soon as a check finds a non-finite energy or
        max |EG| (the device reductions propagate NaN, so a blown-up field is
        never taken for a converged one).
        """
        E = self.compute_energy() if prev_energy is None else float(prev_energy)
        done = 0
        ea = float("inf")
        while done < max_steps:
            n = min(check_every, max_steps - done)
            E, _, ea = self.steps(E, n)
            done += n
            _check_finite(E, ea, done)
            if ea < tol:
                break
        return E, ea, done

    def minimize_fast(self, tol: float = 1e-6, max_steps: int = 200000, chunk: int = 10000, polish: int = 200,
                      form: str = "fp32c", check_every: int = 20, autotune: bool = False):
        """Relax to max |EG| < tol with the FP32 form doing the bulk of the steps.

        The FP32 form (``form``: "fp32c", the compensated one, or "fp32") runs in
        chunks of at most ``chunk`` steps; after each chunk the state is
        converted to FP64 in place and polished for at most ``polish`` steps
        (max |EG| of an FP32 field has a floor of ~1e-4 from the field's own
        rounding, so convergence is judged on the FP64 field).  The next chunk is
        sized from the decay of the polished max |EG| over the last two checks;
        if a chunk improves it by less than 30 % the FP32 form has stalled and
        the relaxation finishes in FP64.  The launch shapes (tile variants and
        fused sweeps) are kept per precision: as set before the call for the
        precision the context is in, the same for the other one, or, with
        ``autotune``, tuned once for each.  The context ends in FP64 on the
        relaxed state.

        Returns ``(energy, err_absmax, fp32_steps, fp64_steps)``.
        """
        if form not in ("fp32c", "fp32"):
            raise ValueError('form must be "fp32c" or "fp32"')

        def shapes():
            if autotune:
                self.autotune_tiles(reps=3)
            return self.tile_variants()

        def switch(prec, info):
            self.set_precision(prec)
            self.set_tile_variants(info["variants"])
            # the compensated form has no 4+1 sweep; every form keeps the paired schedule
            self.set_fused_stages(info["fused23"], info["fused41"] and prec != "fp32c", pairs=info["pairs"])

        first = self.precision
        self.set_precision("fp64")
        tiles64 = shapes()
        self.set_precision(form)
        tiles32 = shapes() if (autotune or first == form) else self.tile_variants()
        if first != "fp64" and not autotune:  # shapes set for the FP32 form: keep them for FP64 too
            tiles64 = tiles32
        E = self.compute_energy()
        n32 = n64 = 0
        checks = []
        size = chunk
        ea = float("inf")
        while True:
            n = min(size, max_steps - n32 - n64)
            if n > 0:
                E, _, ea = self.steps(E, n)
                n32 += n
                _check_finite(E, ea, n32 + n64)
            switch("fp64", tiles64)
            E = self.compute_energy()
            E, ea, npol = self.minimize(tol, polish, check_every, prev_energy=E)
            n64 += npol
            if ea < tol or n32 + n64 >= max_steps:
                return E, ea, n32, n64
            checks.append((n32 + n64, ea))
            if len(checks) >= 2 and checks[-1][1] > 0.7 * checks[-2][1]:
                E, ea, ntail = self.minimize(tol, max(0, max_steps - n32 - n64), max(check_every, 100),
                                             prev_energy=E)
                return E, ea, n32, n64 + ntail
            if len(checks) >= 2 and checks[-1][1] < checks[-2][1]:
                (s0, e0), (s1, e1) = checks[-2], checks[-1]
                need = np.log(tol / e1) / (np.log(e1 / e0) / (s1 - s0))
                size = int(min(max(1.05 * need, chunk / 10), chunk))
            else:
                size = chunk
            switch(form, tiles32)

    # -- observables --------------------------------------------------------
    def compute_energy(self) -> float:
        out = C.c_double()
        self._check(_lib.lib().ss_energy(self._ctx, C.byref(out)))
        return out.value

    def compute_charge(self, kind: int) -> float:
        out = C.c_double()
        self._check(_lib.lib().ss_charge(self._ctx, int(kind), C.byref(out)))
        return out.value

    def observables(self) -> dict:
        """Energy plus every charge the theory's observables module defines.

        core/simulation.py:168-192 (same keys).
        """
        obs = {"energy": self.observables_mod.compute_energy(self)}
        for key, fn in (("skyrmion_number", "compute_skyrmion_number"),
                        ("vortex_number", "compute_vortex_number"),
                        ("magnetic_charge", "compute_magnetic_charge"),
                        ("electric_charge", "compute_electric_charge")):
            f = getattr(self.observables_mod, fn, None)
            if f is not None:
                obs[key] = f(self)
        return obs

    # -- per-site densities (core/simulation.py:194-386) ---------------------
    # The reference fills its shared ``en`` buffer; here each method also
    # returns the values, shape (xlen, ylen), and keeps them in ``self.en``.
    def compute_energy_density(self) -> np.ndarray:
        """Per-site energy times gv (core/simulation.py:194-209)."""
        return self.density(-1)

    def density(self, kind: int) -> np.ndarray:
        """Per-site density times gv: kind -1 energy, >= 0 the theory's charge kinds."""
        out = np.empty(int(self.rp.dim_grid))
        self._check(_lib.lib().ss_density(self._ctx, int(kind), out, out.size))
        self.en = out.reshape(int(self.rp.xlen), int(self.rp.ylen))
        return self.en

    def aux_components(self, kind: int) -> int:
        return int(_lib.lib().ss_aux_components(self._ctx, int(kind)))

    def aux_density(self, kind: int, comp: int = 0) -> np.ndarray:
        """Auxiliary per-site quantity ``kind`` (``_lib.AUX_*``), component ``comp``."""
        out = np.empty(int(self.rp.dim_grid))
        self._check(_lib.lib().ss_aux_density(self._ctx, int(kind), int(comp), out, out.size))
        return out.reshape(int(self.rp.xlen), int(self.rp.ylen))

    def _aux_field(self, kind: int, what: str) -> np.ndarray:
        n = self.aux_components(kind)
        if n == 0:
            raise AttributeError(f"theory {self.theory.THEORY_SPEC.name!r} defines no {what}")
        return np.stack([self.aux_density(kind, c) for c in range(n)])

    def _charge_density(self, key: str) -> np.ndarray:
        kinds = getattr(self.theory.kernels, "CHARGE_KINDS", {})
        if key not in kinds:
            raise AttributeError(f"theory {self.theory.THEORY_SPEC.name!r} defines no {key.replace('_', ' ')}")
        return self.density(kinds[key])

    def compute_higgs_density(self) -> np.ndarray:
        """|psi|^2 per site (core/simulation.py:211-226)."""
        self.en = self._aux_field(_lib.AUX_HIGGS, "Higgs norm")[0]
        return self.en

    def compute_higgs1_density(self) -> np.ndarray:
        """|psi_1|^2 per site (core/simulation.py:228-243)."""
        self.en = self._aux_field(_lib.AUX_HIGGS1, "first Higgs norm")[0]
        return self.en

    def compute_higgs2_density(self) -> np.ndarray:
        """|psi_2|^2 per site (core/simulation.py:245-260)."""
        self.en = self._aux_field(_lib.AUX_HIGGS2, "second Higgs norm")[0]
        return self.en

    def compute_phase_difference_density(self) -> np.ndarray:
        """cos(arg psi_1 - arg psi_2) per site (core/simulation.py:262-277)."""
        self.en = self._aux_field(_lib.AUX_PHASE_DIFFERENCE, "phase difference")[0]
        return self.en

    def compute_magnetic_flux_density(self, which: int) -> np.ndarray:
        """Flux density ``which`` times gv/(2 pi q) per site (core/simulation.py:279-299)."""
        kinds = getattr(self.theory.kernels, "CHARGE_KINDS", {})
        key = {0: "vortex_number_0", 1: "vortex_number_1"}.get(int(which), "vortex_number")
        return self._charge_density(key if key in kinds else "vortex_number")

    def magnetic_field(self) -> np.ndarray:
        """Magnetic flux density field, shape (components, xlen, ylen) (core/simulation.py:301-317)."""
        self.MagneticFluxDensity = self._aux_field(_lib.AUX_MAGNETIC_FIELD, "magnetic field")
        return self.MagneticFluxDensity

    def compute_magnetic_charge_density(self) -> np.ndarray:
        """core/simulation.py:319-334."""
        return self._charge_density("magnetic_charge")

    def compute_electric_charge_density(self) -> np.ndarray:
        """core/simulation.py:336-351."""
        return self._charge_density("electric_charge")

    def compute_noether_charge_density(self) -> np.ndarray:
        """core/simulation.py:353-368."""
        return self._charge_density("noether_charge")

    def supercurrent(self) -> np.ndarray:
        """Supercurrent field, shape (components, xlen, ylen) (core/simulation.py:370-386)."""
        self.Supercurrent = self._aux_field(_lib.AUX_SUPERCURRENT, "supercurrent")
        return self.Supercurrent

    def energy_gradient(self, mode: str = "reference") -> np.ndarray:
        """EG of the current field (the theory's gradient kernel with Temp = Field)."""
        out = np.empty(int(self.rp.dim_fields))
        self._check(_lib.lib().ss_variation(self._ctx, _VAR[mode], out, out.size))
        return out.reshape(self.field_shape)

    def history(self, nsteps: int):
        """(energies, errs) of the last ``nsteps`` steps of the last stepping
        call: energies[k] is the energy after step k, errs[k] the signed max
        of step k's stage-4 energy gradient."""
        hE = np.zeros(int(nsteps))
        hR = np.zeros(int(nsteps))
        self._check(_lib.lib().ss_step_history(self._ctx, hE, hR, int(nsteps)))
        return hE, hR

    def time_passes(self, prev_energy: float, nsteps: int) -> np.ndarray:
        """Mean ms of [pass1, energy finalize, pass2, pass3, pass4, err finalize]."""
        out = np.zeros(6)
        self._check(_lib.lib().ss_time_passes(self._ctx, float(prev_energy), int(nsteps), out))
        return out

    NUM_SHAPES = 9

    def autotune_tiles(self, reps: int = 10) -> dict:
        """Time every fused pass in each shape variant on the live buffers and keep the
        fastest per pass (the relaxation state is unchanged).  Variants 0-2 are LDS
        tiles (default rows, halved, quartered), 3-5 strips of 64/128/256 rows and
        64 columns in the streaming form, 6-8 the same strips 128 columns wide; the
        fused sweeps of passes 2+3 and 4+1 are timed in every variant they run in
        (tiles where two fields fit in LDS, strips).  Returns the chosen variants,
        their rows per tile/strip, the fused flags and the ms tables ([pass][variant],
        fused 2+3 and 4+1 per variant, the paired sweeps 1+2 and 3+4 per variant;
        0 where a sweep does not run)."""
        chosen = np.zeros(4, dtype=np.int32)
        nv = self.NUM_SHAPES
        ms = np.zeros(8 * nv)
        self._check(_lib.lib().ss_autotune_tiles(self._ctx, int(reps), chosen, ms))
        info = self.tile_variants()
        info["ms"] = ms[:4 * nv].reshape(4, nv)
        info["ms_fused23"] = ms[4 * nv:5 * nv]
        info["ms_fused41"] = ms[5 * nv:6 * nv]
        info["ms_pairs12"] = ms[6 * nv:7 * nv]
        info["ms_pairs34"] = ms[7 * nv:]
        return info

    def set_fused_stages(self, fuse23: bool, fuse41: bool = False, pairs: bool = False) -> None:
        """Run passes 2+3 and/or pass 4 with the next step's pass 1 as single sweeps,
        or (``pairs``, alone) the step as two sweeps, passes 1+2 and 3+4 (see
        ``ss_set_fused_stages``)."""
        flags = (_lib.FUSE_23 if fuse23 else 0) | (_lib.FUSE_41 if fuse41 else 0) | (_lib.FUSE_PAIRS if pairs else 0)
        self._check(_lib.lib().ss_set_fused_stages(self._ctx, flags))

    def set_tile_variants(self, variants) -> None:
        """Shape variant (0..8, see ``autotune_tiles``) of passes 1..4."""
        v = np.ascontiguousarray(variants, dtype=np.int32)
        if v.shape != (4,):
            raise ValueError("tile variants: one entry per fused pass (4)")
        self._check(_lib.lib().ss_set_tile_variants(self._ctx, v))

    def tile_variants(self) -> dict:
        v = np.zeros(4, dtype=np.int32)
        rows = np.zeros(4, dtype=np.int32)
        self._check(_lib.lib().ss_get_tile_variants(self._ctx, v, rows))
        on = C.c_int()
        self._check(_lib.lib().ss_get_fused_stages(self._ctx, C.byref(on)))
        return {"variants": v.tolist(), "rows": rows.tolist(), "fused23": bool(on.value & _lib.FUSE_23),
                "fused41": bool(on.value & _lib.FUSE_41), "pairs": bool(on.value & _lib.FUSE_PAIRS),
                "comp12": bool(on.value & _lib.FUSE_COMP12)}

    # -- output bundle and resume (core/simulation.py:411-510, core/io.py) ---
    def save_output(self, output_dir: str | None = None, precision: int = 32) -> str:
        """Write the theory's output bundle (Field.dat, lattice files, densities)
        in the reference's text format; returns the directory."""
        import inspect
        from pathlib import Path
        if output_dir is None:
            output_dir = str(Path(self.theory.io.__file__).resolve().parent / "results")
        Path(output_dir).mkdir(parents=True, exist_ok=True)
        kinds = getattr(self.theory.kernels, "CHARGE_KINDS", {})

        def dens(key):
            return self.density(kinds[key]).ravel() if key in kinds else None

        def aux(kind):
            n = self.aux_components(kind)
            return np.stack([self.aux_density(kind, c) for c in range(n)]).ravel() if n else None

        kwargs = dict(output_dir=output_dir, h_Field=self.get_field().ravel(),
                      h_EnergyDensity=self.compute_energy_density().ravel(),
                      h_MagneticFluxDensity=aux(_lib.AUX_MAGNETIC_FIELD),
                      h_BaryonDensity=dens("skyrmion_number"),
                      h_VortexDensity=dens("vortex_number"),
                      h_Supercurrent=aux(_lib.AUX_SUPERCURRENT),
                      h_HiggsDensity=aux(_lib.AUX_HIGGS),
                      h_MagneticChargeDensity=dens("magnetic_charge"),
                      h_ElectricChargeDensity=dens("electric_charge"),
                      h_NoetherChargeDensity=dens("noether_charge"),
                      h_Higgs1Density=aux(_lib.AUX_HIGGS1),
                      h_Higgs2Density=aux(_lib.AUX_HIGGS2),
                      h_grid=self.create_grid().ravel(),
                      xlen=int(self.p_i_h[0]), ylen=int(self.p_i_h[1]),
                      number_coordinates=int(self.p_i_h[3]), number_total_fields=int(self.p_i_h[4]),
                      precision=precision)
        accepted = set(inspect.signature(self.theory.io.output_data_bundle).parameters)
        self.theory.io.output_data_bundle(**{k: v for k, v in kwargs.items() if k in accepted})
        return output_dir

    def load_output(self, output_dir: str, velocity_zero: bool = True) -> None:
        """Resume from a bundle's Field.dat (velocity set to zero)."""
        import os
        from .io import read_field_dat
        F = read_field_dat(os.path.join(output_dir, "Field.dat"), int(self.p_i_h[0]), int(self.p_i_h[1]),
                           int(self.p_i_h[4]))
        self.set_field(F)
        if velocity_zero:
            self.set_velocity(np.zeros_like(F))

    def synchronize(self) -> None:
        self._check(_lib.lib().ss_synchronize(self._ctx))

    def memory_bytes(self) -> int:
        n = C.c_size_t()
        self._check(_lib.lib().ss_memory_bytes(self._ctx, C.byref(n)))
        return int(n.value)

    def print_instructions(self) -> None:
        print("Headless relaxation core: drive it with sim.step(E) / sim.steps(E, n) / sim.minimize().")
