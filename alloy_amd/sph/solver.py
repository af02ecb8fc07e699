"""SolverDFSPH: the DFSPH liquid plug-in, host side, driving the gfx950 kernels of libdfsph_hip.so.

Same construction, methods, errors and step order as alloy/solvers/sph/solver.py:57-484: ``step`` runs neighbours
-> boundary slots -> density and factor -> divergence solve -> gravity, external forces and XSPH -> CSF surface
tension -> implicit viscosity -> density solve -> shape meters -> advection and wall projection -> sinks, escape box
and box emitters -> speed statistics, all on the current torch stream. Every work buffer is allocated once at
construction; ``step`` allocates nothing.

The mass ledger is int64 fixed point on the device (2^40 per kg, runtime.py:93-153), so active + out == initial + in
holds exactly. ``reneighbor_skin`` reuses superset neighbour rows between rebuilds (neighbors_skin.py; the rows stay
equal to a plain build, bit for bit). Subsystems of the reference that this library does not implement (ambient
pressure, Akinci 2013 tension, soft walls, bodies attached to shapes and the body-attached patch emitters) are
refused with an error naming them, rather than silently ignored.
"""

from __future__ import annotations

import ctypes
import dataclasses
import logging
from collections.abc import Mapping
from typing import ClassVar

import numpy as np
import torch

from . import _abi, csf
from .emitters import EmitterManager
from .config import (PHASE_LIQUID, DFSPHConfig, register_custom_attributes, resolve_viscosity, snap_down,
                     stable_dt)
from .model import Model, ModelFlags, ParticleFlags, State
from .types import SPHReport

log = logging.getLogger("alloy.sph")

PARTICLE_Q, PARTICLE_QD = 1 << 4, 1 << 5  # Newton's StateFlags bits of the coupling notifications

CapacityError = _abi.CapacityError


def hydrostatic_guess(model: Model, state: State, spacing: float) -> np.ndarray:
    """k_i = |g| (z_top - z_i + r) per (world, horizontal 2r bin) column of active liquid, along gravity
    (prime.py:20-50): the initial pressure ``prime`` refines [m^2/s^2]."""
    n = model.particle_count
    out = np.zeros(max(n, 1), dtype=np.float32)
    if n == 0:
        return out
    q = state.particle_q.detach().cpu().numpy().astype(np.float64)
    act = model.sph.phase.cpu().numpy() >= 1
    act &= (model.particle_flags.cpu().numpy() & int(ParticleFlags.ACTIVE)) != 0
    world = model.particle_world.cpu().numpy()
    grav = model.gravity.cpu().numpy().astype(np.float64)
    for w in np.unique(world[act]):
        g = grav[w]
        gn = float(np.linalg.norm(g))
        if gn == 0.0:
            continue
        up = -g / gn
        sel = np.nonzero(act & (world == w))[0]
        z = q[sel] @ up
        e1 = np.cross(up, [1.0, 0.0, 0.0] if abs(up[0]) < 0.9 else [0.0, 1.0, 0.0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(up, e1)
        key = np.stack([np.floor(q[sel] @ e1 / spacing), np.floor(q[sel] @ e2 / spacing)], 1).astype(np.int64)
        _, inv = np.unique(key, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        top = np.full(inv.max() + 1, -np.inf)
        np.maximum.at(top, inv, z)
        out[sel] = (gn * np.maximum(top[inv] - z + 0.5 * spacing, 0.0)).astype(np.float32)
    return out


class ParticleWallReactions:
    """The liquid's reaction on the particle walls (particle_walls.py:236-282 world_work / particle_impulse /
    new_window / harvest): each particle-wall slot's impulse is split over the wall particles in int64 at 2^50 every
    step, summed over an exchange window; ``count`` wall particles in ascending id order."""

    def __init__(self, solver):
        self._s = solver
        self.vid = np.nonzero(solver.tables.wall_volume > 0.0)[0]
        self.count = int(len(self.vid))

    def world_work(self) -> np.ndarray:
        """The last step's work [J] of the walls on the liquid, per world."""
        if not self.count:
            return np.zeros(0)
        W = self._s.model.world_count
        return self._s.buffer(_abi.BUF_PW_WORK, np.int64, (W,)).astype(np.float64) / _abi.WORK_SCALE

    def particle_impulse(self) -> np.ndarray:
        """The exchange window's impulse [N s] per wall particle, (count, 3)."""
        if not self.count:
            return np.zeros((0, 3))
        return self._s.buffer(_abi.BUF_PW_EXCH, np.int64, (self.count, 3)).astype(np.float64) / _abi.IMP_SCALE

    def new_window(self) -> None:
        if self.count:
            self._s._write(_abi.BUF_PW_EXCH, np.zeros((self.count, 3), np.int64))

    def harvest(self, local_to_global, out: torch.Tensor, dt: float) -> None:
        """out[local_to_global[vid]] = the window's impulse / dt (a force [N]) for the wall particles mapped >= 0."""
        if not self.count:
            return
        l2g = np.asarray(local_to_global)[self.vid]
        keep = l2g >= 0
        f = torch.as_tensor(self.particle_impulse()[keep] / float(dt), dtype=out.dtype, device=out.device)
        out[torch.as_tensor(l2g[keep], device=out.device)] = f


def _stream_ptr(device: torch.device):
    if device.type != "cuda":
        return None
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def as_state_tensor(a, name: str, device: torch.device, n: int | None = None):
    """A zero-copy torch view of a state array the caller owns: a torch tensor as it is, otherwise any device array
    exporting DLPack (``__dlpack__``, as Warp arrays and the vendored Newton State's arrays do) or
    ``__cuda_array_interface__``. vec3 arrays come as [n, 3] float32 (a flat [3 n] export is viewed as [n, 3]).
    Host arrays are refused: the boundary takes device memory, and the step writes the caller's q and qd in place."""
    if a is None or isinstance(a, torch.Tensor):
        return a
    if hasattr(a, "__dlpack__"):
        t = torch.from_dlpack(a)
    elif hasattr(a, "__cuda_array_interface__"):
        t = torch.as_tensor(a, device=device)
    else:
        raise ValueError(f"SolverDFSPH: state.{name} is a {type(a).__name__}, neither a torch tensor nor a device array "
                         "exporting __dlpack__ or __cuda_array_interface__")
    want = torch.device(device)
    if t.device.type != want.type or (want.index is not None and t.device.index != want.index):
        raise ValueError(f"SolverDFSPH: state.{name} lives on {t.device}, the solver on {device}")
    if t.dim() == 1 and (n is None or t.numel() == 3 * n) and t.numel() % 3 == 0:
        t = t.view(-1, 3)
    return t


class _StateView:
    """The caller's state seen through torch views of its arrays (as_state_tensor); writes reach the caller's
    memory."""

    __slots__ = ("particle_q", "particle_qd", "particle_f")

    def __init__(self, state, device: torch.device, n: int):
        self.particle_q = as_state_tensor(state.particle_q, "particle_q", device, n)
        self.particle_qd = as_state_tensor(state.particle_qd, "particle_qd", device, n)
        self.particle_f = as_state_tensor(getattr(state, "particle_f", None), "particle_f", device, n)


def device_state(state, device: torch.device, n: int):
    """``state`` itself when its arrays are torch tensors, else a _StateView of them."""
    if isinstance(state.particle_q, torch.Tensor) and isinstance(state.particle_qd, torch.Tensor) and isinstance(
            getattr(state, "particle_f", None), (torch.Tensor, type(None))):
        return state
    return _StateView(state, device, n)


class SolverDFSPH:
    """Divergence-free SPH (Bender & Koschier 2017) with SDF and particle walls, on MI355X."""

    plugin_name: ClassVar[str] = "sph"
    routes: ClassVar[frozenset[str]] = frozenset({"liquid_sph"})
    determinism: ClassVar[str] = "native"
    alloy_clock_safety: ClassVar[float] = 1.0
    Config = DFSPHConfig
    register_custom_attributes = staticmethod(register_custom_attributes)

    def __init__(self, model: Model, config: DFSPHConfig | None = None, *, capacity_per_world: int | None = None):
        self.model = model
        self.config = cfg = config or DFSPHConfig()
        cfg.validate()
        self._refuse_unbuilt(cfg)
        if model.device.type != "cuda":
            raise RuntimeError("SolverDFSPH runs on a HIP device (the model is on "
                               f"{model.device}); there is no CPU path")
        self.device = model.device
        self._check_layout(model)
        self._check_capacity(capacity_per_world)
        self.viscosity = resolve_viscosity(cfg, float(model.sph.viscosity.max().item()) if model.particle_count else 0.0)
        if cfg.boundary_viscosity and self.viscosity != "weiler2018":  # solver.py:164-165
            raise ValueError("SolverDFSPH: boundary_viscosity needs the weiler2018 viscosity "
                             f"(resolved {self.viscosity!r})")
        csf.reject_unread_inputs(model, cfg)  # layout.py:80-95
        self.tension = cfg.surface_tension  # solver.py:83-87, csf.py:298-309
        self.capillary_dt_violations = 0
        self._dt_limit = csf.capillary_limit(model, cfg) if self.tension == "csf" else float("inf")
        if self._dt_limit < cfg.dt_max:  # the capillary step bounds the stable step
            self.config = cfg = dataclasses.replace(cfg, dt_max=max(self._dt_limit, cfg.dt_min))
        self.tension_rh_range = None
        if self.tension == "csf":
            viscous = self.viscosity == "weiler2018"
            self.tension_rh_range = csf.RH_RANGE_VISCOUS if viscous else csf.RH_RANGE_XSPH
            if not viscous and model.particle_count:
                log.warning("SPH csf: low-viscosity liquid (viscosity 'xsph'): a free drop's l = 2 oscillation grows "
                            "instead of decaying; quantitative for the shape, the Laplace pressure and the first "
                            "oscillation periods at R/h %.1f-%.1f only", *csf.RH_RANGE_XSPH)
        phase = model.sph.phase.cpu().numpy()
        mass = model.particle_mass.cpu().numpy()[phase >= 0]
        rho0 = model.sph.rest_density.cpu().numpy()[phase >= 0]
        spread = lambda a: float(np.ptp(a)) / max(float(np.max(np.abs(a))), 1e-30) if len(a) else 0.0  # noqa: E731
        self.pressure_solver = cfg.pressure_solver
        if self.pressure_solver == "cg" and (spread(mass.astype(np.float64)) > 1e-6 or
                                             spread(rho0.astype(np.float64)) > 1e-6):
            log.warning("SolverDFSPH: pressure_solver 'cg' needs one particle mass and rest density; using 'jacobi'")
            self.pressure_solver = "jacobi"
        self.lib = _abi.load_library()
        self._h = None
        self._warm_model = None  # learned warm-start parameters attached to the native handle, if any
        self._create_native()
        self._flags0 = model.particle_flags.clone()
        self._phase0 = model.sph.phase.clone()
        W = model.world_count
        self.emitters = EmitterManager(model, cfg, self._upload_emitters, self._emit_cursors, eager=False)
        self.pw = ParticleWallReactions(self)
        self._soft_back = 0.0     # after a staggered sync: move the particle walls back by v dt once (coupling)
        self._pass_saved = None   # the exchange's solver state, restored by a restart pass
        if self.emitters.active:  # the escape box
            self.emitters.flush(force=True)
        log.info("SolverDFSPH: %d particles in %d worlds, r=%g m, h=%g m, %d shapes, %d wall particles, solver=%s",
                 model.particle_count, W, cfg.particle_radius, cfg.h, model.shape_count, self.tables.wall_count,
                 self.pressure_solver)

    def _upload_emitters(self, st) -> None:
        self._call(self.lib.dfsph_set_emitters, ctypes.byref(st.struct), what="SolverDFSPH emitters")

    def _emit_cursors(self) -> np.ndarray:
        E = len(self.emitters.table.world) if hasattr(self, "emitters") else 0
        if E == 0 or self.emitters.struct is None or self.emitters.struct.struct.emitter_count == 0:
            return np.zeros(E, np.int32)
        n_dev = int(self.emitters.struct.struct.emitter_count)  # the table on the device (a newer one may wait)
        out = np.zeros(E, np.int32)
        dev = np.empty(n_dev, np.int32)
        self._call(self.lib.dfsph_read_buffer, _abi.BUF_EMIT_CURSOR, dev.ctypes.data_as(ctypes.c_void_p), dev.nbytes)
        out[:n_dev] = dev
        return out

    # ------------------------------------------------------------------------------------------------- checks
    def _create_native(self) -> None:
        """The native handle for the model's current shapes and layout (solver.py:127-168 allocates the same)."""
        model, cfg = self.model, self.config
        self.tables = _abi.layout_from_model(model, cfg)
        self._params = _abi.params_from_config(cfg, self.tables.thinnest)
        self._params.pressure_solver = 0 if self.pressure_solver == "cg" else 1
        self._params.xsph_c = csf.xsph_coefficient(cfg, self.viscosity)  # solver.py:87, 299
        self._params.viscosity_implicit = int(self.viscosity == "weiler2018")
        self._mu = model.sph.viscosity.to(torch.float32).contiguous()
        self._wall_volume = torch.tensor(self.tables.wall_volume, device=self.device)
        self._gravity = model.gravity.reshape(-1).contiguous().clone()
        m = _abi.DfsphModel()
        m.mass, m.rest_density = _dptr(model.particle_mass), _dptr(model.sph.rest_density)
        m.flags, m.phase, m.world = _dptr(model.particle_flags), _dptr(model.sph.phase), _dptr(model.particle_world)
        m.gravity, m.wall_volume = _dptr(self._gravity), _dptr(self._wall_volume)
        m.viscosity = _dptr(self._mu)
        if self.tension == "csf":
            self._sigma = model.sph.surface_tension.to(torch.float32).contiguous()
            self._gamma = torch.tensor(csf.liquid_gamma(model, cfg), device=self.device)
            m.surface_tension, m.csf_gamma = _dptr(self._sigma), _dptr(self._gamma)
        self._model_struct = m
        layout = self.tables.struct(model.particle_count)
        h = ctypes.c_void_p()
        _abi.check(self.lib.dfsph_create(ctypes.byref(self._params), ctypes.byref(layout), ctypes.byref(m),
                                         ctypes.byref(h)), self.lib, "SolverDFSPH")
        self._h = h

    @staticmethod
    def _refuse_unbuilt(cfg: DFSPHConfig) -> None:
        unbuilt = [
            (cfg.surface_tension not in ("none", "csf"), f"surface_tension={cfg.surface_tension!r}"),
            (cfg.ambient_pressure is not None, "ambient_pressure"),
        ]
        for bad, name in unbuilt:
            if bad:
                raise ValueError(f"SolverDFSPH: {name} is not built in this DFSPH implementation")

    @staticmethod
    def _check_layout(model: Model) -> None:
        """layout.py:47-72: SPH particles in a world, worlds contiguous."""
        if model.particle_count == 0:
            return
        world = model.particle_world.cpu().numpy()
        phase = model.sph.phase.cpu().numpy()
        if np.any(world[phase >= 0] < 0):
            raise ValueError("SPH particles must belong to a world (particle_world >= 0)")
        if np.any(np.diff(world) < 0):
            raise ValueError("particles must be contiguous per world (Model.particle_world sorted)")

    def _check_capacity(self, capacity_per_world: int | None) -> None:
        if capacity_per_world is None:
            return
        ws = self.model.world_start_np
        phase = self.model.sph.phase.cpu().numpy()
        for w in range(self.model.world_count):
            have = int(np.count_nonzero(phase[ws[w]:ws[w + 1]] >= 0))
            if have < capacity_per_world:
                raise CapacityError(f"SolverDFSPH: world {w} has {have} SPH particle slots, "
                                    f"capacity_per_world={capacity_per_world}")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self.lib.dfsph_destroy(h)
            self._h = None

    # ------------------------------------------------------------------------------------------------- step
    def _call(self, fn, *args, what: str = "SolverDFSPH") -> None:
        _abi.check(fn(self._h, *args), self.lib, what)

    @staticmethod
    def _checked_state(state: State, n: int) -> None:
        for name in ("particle_q", "particle_qd"):
            t = getattr(state, name)
            if t.dtype != torch.float32 or not t.is_contiguous() or tuple(t.shape) != (n, 3):
                raise ValueError(f"SolverDFSPH: state.{name} must be a contiguous float32 [{n}, 3] tensor")

    def step(self, state_in: State, state_out: State, control=None, contacts=None, dt: float = 0.0, *,
             observables=None) -> None:
        """One DFSPH step from ``state_in`` into ``state_out`` (the same object allowed), solver.py:218-268."""
        del control, contacts, observables
        if not dt > 0.0:
            raise ValueError(f"SolverDFSPH.step: dt must be > 0, got {dt}")
        n = self.model.particle_count
        if n == 0:
            return
        same = state_out is state_in
        state_in = device_state(state_in, self.device, n)
        state_out = state_in if same else device_state(state_out, self.device, n)
        self._checked_state(state_in, n)
        self._checked_state(state_out, n)
        if state_out is not state_in:
            state_out.particle_q.copy_(state_in.particle_q)
            state_out.particle_qd.copy_(state_in.particle_qd)
        if self.emitters.dirty:
            self.emitters.flush()
        if self.tension == "csf" and dt > self._dt_limit * (1.0 + 1.0e-6):  # tension.py:280-287: never silent
            if self.capillary_dt_violations == 0:
                log.warning("SPH csf: dt %.3g s is above the capillary step limit %.3g s; the tension laws do not "
                            "cover it", dt, self._dt_limit)
            self.capillary_dt_violations += 1
        if self._soft_back > 0.0 and self.pw.count:  # coupling_hooks.py:170-180, before the grid
            vid = torch.as_tensor(self.pw.vid, device=self.device)
            q, qd = state_out.particle_q, state_out.particle_qd
            q[vid] = q[vid] + (-self._soft_back) * qd[vid]
        self._soft_back = 0.0
        f = getattr(state_in, "particle_f", None)
        self._call(self.lib.dfsph_step, _dptr(state_out.particle_q), _dptr(state_out.particle_qd), _dptr(f),
                   float(dt), _stream_ptr(self.device), what="SolverDFSPH.step")

    def set_check_interval(self, every: int) -> None:
        """Iterations between host convergence checks of the pressure solves (autotuned; results unchanged)."""
        self._call(self.lib.dfsph_set_check_interval, int(every))

    def set_option(self, option: int, value: int) -> None:
        """A tuning option of the native step (``_abi.OPT_*``); results are bit-identical for every value."""
        self._call(self.lib.dfsph_set_option, int(option), int(value))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int32()
        self._call(self.lib.dfsph_get_option, int(option), ctypes.byref(v))
        return int(v.value)

    def set_graphs(self, on: bool) -> None:
        """Replay the step's segments as captured hipGraphs (default) or launch every kernel eagerly."""
        self._call(self.lib.dfsph_set_graphs, int(bool(on)))

    def set_warm_start(self, kappa_div: torch.Tensor | None, kappa_den: torch.Tensor | None) -> None:
        """Learned warm start: initial kappa [n] of each solve (None: the reference rule). Tensors must stay alive
        while set."""
        self._warm = (kappa_div, kappa_den)
        self._call(self.lib.dfsph_set_warm_start_kappa, _dptr(kappa_div), _dptr(kappa_den))

    # ------------------------------------------------------------------------------------------------- reports
    def buffer(self, which: int, dtype, shape) -> np.ndarray:
        """Host copy of a solver buffer (tests, diagnostics; syncs)."""
        if self.emitters.dirty:
            self.emitters.flush()
        out = np.empty(shape, dtype=dtype)
        self._call(self.lib.dfsph_read_buffer, which, out.ctypes.data_as(ctypes.c_void_p), out.nbytes)
        return out

    def _raw_report(self) -> dict[str, np.ndarray]:
        W = self.model.world_count
        arrs = {k: np.zeros(W, dt) for k, dt in (
            ("density_err", np.float32), ("divergence_err", np.float32), ("iters_density", np.int32),
            ("iters_divergence", np.int32), ("neighbor_overflow", np.int32), ("max_speed", np.float32),
            ("density_err_max", np.float32), ("density_tol", np.float32), ("projected", np.int32),
            ("swept", np.int32), ("nonfinite", np.int32), ("n_active", np.float64))}
        rep = _abi.DfsphReport()
        ct = {np.float32: ctypes.c_float, np.int32: ctypes.c_int32, np.float64: ctypes.c_double}
        for k, a in arrs.items():
            setattr(rep, k, a.ctypes.data_as(ctypes.POINTER(ct[a.dtype.type])))
        self._call(self.lib.dfsph_report, ctypes.byref(rep), _stream_ptr(self.device))
        return arrs

    def report(self) -> SPHReport:
        """Per-world report of the last step (runtime.py:136-147); the step's sync point."""
        r = self._raw_report()
        mi = self.config.max_iterations
        led = self.mass_ledger()
        kg = lambda a: a.astype(np.float64) / _abi.FIX_SCALE  # noqa: E731
        W = self.model.world_count
        return SPHReport(
            density_err=r["density_err"].astype(np.float64), divergence_err=r["divergence_err"].astype(np.float64),
            iters_density=r["iters_density"], iters_divergence=r["iters_divergence"], mass_active=kg(led["active"]),
            mass_in=kg(led["in"]), mass_out=kg(led["out"]), escaped=self.buffer(_abi.BUF_ESCAPED, np.int32, (W,)),
            neighbor_overflow=r["neighbor_overflow"], max_speed=r["max_speed"].astype(np.float64),
            unconverged=(r["iters_density"] >= mi) | (r["iters_divergence"] >= mi),
            density_tol=r["density_tol"].astype(np.float64), tension=self.tension,
            tension_qualitative=False, tension_rh_range=self.tension_rh_range,
            capillary_dt_violations=self.capillary_dt_violations)

    def diagnostics(self) -> dict[str, np.ndarray]:
        """Extra per-world counters (runtime.py:155-163)."""
        r = self._raw_report()
        W = self.model.world_count
        return {"density_err_max": r["density_err_max"].astype(np.float64), "projected": r["projected"],
                "nonfinite": r["nonfinite"], "swept": r["swept"], "n_active": r["n_active"],
                "emitter_overflow": self.buffer(_abi.BUF_EMIT_OVERFLOW, np.int32, (W,))}

    def mass_ledger(self) -> dict[str, np.ndarray]:
        """Per-world mass ledger, int64 at ``scale`` per kg (runtime.py:136-140): active + out == initial + in,
        exactly."""
        W = self.model.world_count
        self._call(self.lib.dfsph_active_mass, _stream_ptr(self.device))
        return {"initial": self.buffer(_abi.BUF_MASS_INITIAL, np.int64, (W,)),
                "in": self.buffer(_abi.BUF_MASS_IN, np.int64, (W,)),
                "out": self.buffer(_abi.BUF_MASS_OUT, np.int64, (W,)),
                "active": self.buffer(_abi.BUF_MASS_ACTIVE, np.int64, (W,)), "scale": np.float64(_abi.FIX_SCALE)}

    def shape_impulse(self) -> np.ndarray:
        """Cumulative fluid impulse on each shape since the last reset [N s], (S, 3) float64 (solver.py:399-401)."""
        S = max(self.model.shape_count, 1)
        raw = self.buffer(_abi.BUF_SHAPE_METER, np.int64, (S, 3))
        return raw.astype(np.float64) / float(2**50)

    def reset_shape_meters(self) -> None:
        self._call(self.lib.dfsph_zero_stats, _stream_ptr(self.device))

    # ------------------------------------------------------------------------------------------------- helpers
    def prime(self, state: State, dt: float, tol: float = 1.0e-6, max_iterations: int = 2000) -> int:
        """Give a fluid at rest its hydrostatic pressure before the first step (prime.py:53-115); returns the
        iterations used."""
        if self.model.particle_count == 0:
            return 0
        state = device_state(state, self.device, self.model.particle_count)
        guess = torch.tensor(hydrostatic_guess(self.model, state, self.config.spacing), device=self.device)
        it = ctypes.c_int32(0)
        self._call(self.lib.dfsph_prime, _dptr(state.particle_q), _dptr(state.particle_qd), _dptr(guess), float(dt),
                   float(tol), int(max_iterations), ctypes.byref(it), _stream_ptr(self.device), what="prime")
        return int(it.value)

    def density_ratio(self, state: State) -> tuple[np.ndarray, np.ndarray]:
        """Density ratio of every particle and the active-liquid mask, from fresh lists and slots (solver.py:386-393)."""
        n = self.model.particle_count
        if n == 0:
            return np.zeros(0), np.zeros(0, dtype=bool)
        state = device_state(state, self.device, n)
        st = _stream_ptr(self.device)
        self._call(self.lib.dfsph_build_neighbors, _dptr(state.particle_q), st)
        self._call(self.lib.dfsph_boundary_slots, _dptr(state.particle_q), _dptr(state.particle_qd), st)
        self._call(self.lib.dfsph_density_alpha, _dptr(state.particle_q), st)
        ratio = self.buffer(_abi.BUF_RATIO, np.float32, (n,))
        active = self.buffer(_abi.BUF_ACTIVE, np.float32, (n,))
        return ratio.copy(), active > 0.0

    def reset(self, state: State, world_mask: np.ndarray | None = None, flags: int | None = None) -> None:
        """Masked worlds back to their build-time flags and phase, warm starts cleared (solver.py:353-362)."""
        del state, flags
        m = self.model
        W = m.world_count
        mask = np.ones(W, bool) if world_mask is None else np.asarray(world_mask, bool)[:W]
        sel = torch.tensor(mask, device=self.device)[m.particle_world.long().clamp(min=0)] & (m.particle_world >= 0)
        m.particle_flags[sel] = self._flags0[sel]
        m.sph.phase[sel] = self._phase0[sel]
        n = m.particle_count
        for which in (_abi.BUF_KV, _abi.BUF_KD):
            k = self.buffer(which, np.float32, (max(n, 1),))
            k[:n][sel.cpu().numpy()] = 0.0
            self._write(which, k)
        # the ledger restarts with the restored particles; the initial mass stays (runtime.py:190-194)
        for which, dt in ((_abi.BUF_ESCAPED, np.int32), (_abi.BUF_EMIT_OVERFLOW, np.int32), (_abi.BUF_MASS_IN, np.int64),
                          (_abi.BUF_MASS_OUT, np.int64), (_abi.BUF_FLUX_E, np.int64), (_abi.BUF_FLUX_G, np.int64)):
            a = self.buffer(which, dt, (W,))
            a[mask] = 0
            self._write(which, a)
        picked = self.emitters.rewound_emitters(mask)  # emitters.rewind: clock and cursor back to 0
        E = len(self.emitters.table.world)
        if E and len(picked):
            for which, dt in ((_abi.BUF_EMIT_T, np.float32), (_abi.BUF_EMIT_CURSOR, np.int32)):
                a = self.buffer(which, dt, (E,))
                a[picked] = 0
                self._write(which, a)

    def _write(self, which: int, a: np.ndarray) -> None:
        if self.emitters.dirty:
            self.emitters.flush()
        a = np.ascontiguousarray(a)
        self._call(self.lib.dfsph_write_buffer, which, a.ctypes.data_as(ctypes.c_void_p), a.nbytes)

    def choose_dt(self, control_dt: float) -> float:
        """CFL step from the last step's max speed, snapped onto the D4 grid (solver.py:424-426)."""
        v = float(np.max(self._raw_report()["max_speed"])) if self.model.world_count else 0.0
        return snap_down(stable_dt(self.config, v), control_dt)

    def alloy_stable_dt(self, state: State, contacts, out: torch.Tensor) -> None:
        """Per-world CFL step from the state's liquid speeds (runtime.py _stable_dt_kernel)."""
        del contacts
        m, cfg = self.model, self.config
        fl = ((m.particle_flags & int(ParticleFlags.ACTIVE)) != 0) & (m.sph.phase >= PHASE_LIQUID)
        sp = torch.linalg.vector_norm(state.particle_qd, dim=1)
        sp = torch.where(torch.isfinite(sp), sp, torch.full_like(sp, 1.0e30))
        vmax = torch.zeros(m.world_count, device=self.device).scatter_reduce(
            0, m.particle_world.long().clamp(min=0), torch.where(fl, sp, torch.zeros_like(sp)), "amax")
        dt = torch.clamp(cfg.cfl * cfg.spacing / vmax.clamp(min=1e-30), cfg.dt_min, cfg.dt_max)
        out.copy_(torch.where(vmax > 0, dt, torch.full_like(dt, cfg.dt_max)).to(out.dtype))

    def alloy_energy(self, state: State, z_ref: torch.Tensor, out: torch.Tensor) -> None:
        """KE + gravity PE about (0, 0, z_ref[w]) per world [J] (runtime.py _energy_kernel)."""
        m = self.model
        fl = ((m.particle_flags & int(ParticleFlags.ACTIVE)) != 0) & (m.sph.phase >= PHASE_LIQUID)
        w = m.particle_world.long().clamp(min=0)
        v = state.particle_qd.double()
        x = state.particle_q.double().clone()
        x[:, 2] -= z_ref.double()[w]
        g = m.gravity.double()[w]
        e = m.particle_mass.double() * (0.5 * (v * v).sum(1) - (g * x).sum(1))
        e = torch.where(fl, e, torch.zeros_like(e))
        out.copy_(torch.zeros(m.world_count, dtype=torch.float64, device=self.device).index_add_(0, w, e)
                  .to(out.dtype))

    def alloy_capacity_report(self) -> Mapping[str, int]:
        W = self.model.world_count
        return {"neighbor_overflow": int(self._raw_report()["neighbor_overflow"].sum()),
                "emitter_overflow": int(self.buffer(_abi.BUF_EMIT_OVERFLOW, np.int32, (W,)).sum())}

    def _ledger_keys(self):
        out = [("sph.mass_in", _abi.BUF_MASS_IN, np.int64), ("sph.mass_out", _abi.BUF_MASS_OUT, np.int64),
               ("sph.escaped", _abi.BUF_ESCAPED, np.int32)]
        if self.emitters.table.world:
            out += [("sph.emit_t", _abi.BUF_EMIT_T, np.float32), ("sph.emit_cursor", _abi.BUF_EMIT_CURSOR, np.int32)]
        return out

    def alloy_state_arrays(self, state: State | None = None) -> dict[str, np.ndarray]:
        """Arrays a snapshot must restore bitwise for this solver to replay (coupling_hooks.py:115-126,
        runtime.py:178-179 and 389-393: the ledger and the emitters' clocks and cursors)."""
        n = max(self.model.particle_count, 1)
        out = {"sph.kappa_v": self.buffer(_abi.BUF_KV, np.float32, (n,)),
               "sph.kappa_d": self.buffer(_abi.BUF_KD, np.float32, (n,)),
               "sph.phase": self.model.sph.phase.cpu().numpy().copy(),
               "sph.flags": self.model.particle_flags.cpu().numpy().copy()}
        if self.model.shape_count:  # a re-run period must not count its walls' impulse twice
            out["sph.meter_j"] = self.buffer(_abi.BUF_SHAPE_METER, np.int64, (self.model.shape_count, 3))
        W, E = self.model.world_count, len(self.emitters.table.world)
        for key, which, dt in self._ledger_keys():
            out[key] = self.buffer(which, dt, (E if key.startswith("sph.emit") else W,))
        return out

    # ------------------------------------------------------------------ coupling hooks (coupling_hooks.py:37-142)
    def _pass_arrays(self) -> dict[str, np.ndarray]:
        """What every pass of one coupling exchange starts from: the warm starts, the shape meters and, with any
        emitter, sink or escape box, what the emitter pass mutates (runtime.py:364-386)."""
        n = max(self.model.particle_count, 1)
        out = {"sph.kappa_v": self.buffer(_abi.BUF_KV, np.float32, (n,)),
               "sph.kappa_d": self.buffer(_abi.BUF_KD, np.float32, (n,))}
        if self.model.shape_count:
            out["sph.meter_j"] = self.buffer(_abi.BUF_SHAPE_METER, np.int64, (self.model.shape_count, 3))
        if self.emitters.active:
            out["sph.phase"] = self.model.sph.phase.cpu().numpy().copy()
            out["sph.flags"] = self.model.particle_flags.cpu().numpy().copy()
            W, E = self.model.world_count, len(self.emitters.table.world)
            for key, which, dt in self._ledger_keys():
                out[key] = self.buffer(which, dt, (E if key.startswith("sph.emit") else W,))
        return out

    def coupling_notify_input_state_update(self, state: State, flags: int, *, iteration_restart: bool = False,
                                           dt: float = 0.0) -> None:
        """A new coupling pass (coupling_hooks.py:43-70): a restart pass restores the exchange's warm starts, shape
        meters and emitter state, so every pass starts from the same solver state and the meters count the pass
        that sticks; any other notification saves them. Either way a new particle-wall impulse window begins."""
        del state
        if self.model.particle_count == 0:
            return
        if int(flags) & PARTICLE_Q and self.pw.count:
            self._soft_back = float(dt)
        if iteration_restart:
            if self._pass_saved is not None:
                self.alloy_restore_state_arrays(self._pass_saved)
        else:
            self._pass_saved = self._pass_arrays()
        self.pw.new_window()

    def coupling_harvest_proxy_wrenches(self, body_local_to_proxy_global, out_body_f, *, body_qd_before=None,
                                        state=None, state_out=None, contacts=None, dt: float = 0.0) -> None:
        """Body proxies (coupling_hooks.py:72-85): this library has no bodies, so there is nothing to harvest."""
        if len(body_local_to_proxy_global) == 0:
            return
        raise NotImplementedError("SolverDFSPH: shapes attached to bodies are not built in this DFSPH implementation")

    def coupling_rewind_proxy_particle(self, particle_local_to_proxy_global, state, coupling_forces,
                                       particle_gravity_acceleration, dt: float) -> None:
        """Lagged particle proxies are refused for particle walls (coupling_hooks.py:87-94)."""
        if len(particle_local_to_proxy_global) and self.pw.count:
            raise ValueError("SolverDFSPH particle walls need a staggered proxy pair (PairPolicy / Proxy mode "
                             "'staggered' with Aitken relaxation); lagged proxies diverge for a liquid-wall pair")

    def coupling_harvest_proxy_particle_forces(self, particle_local_to_proxy_global, out_particle_f, *,
                                               dt: float = 0.0, **kwargs) -> None:
        """Each mapped wall particle's exchange-averaged force from the liquid (coupling_hooks.py:96-110)."""
        del kwargs
        if len(particle_local_to_proxy_global) == 0:
            return
        if not dt > 0.0:
            raise ValueError("SolverDFSPH harvest requires dt > 0")
        if self.pw.count == 0:
            raise ValueError("SolverDFSPH: proxy particles but no particle walls")
        self.pw.harvest(particle_local_to_proxy_global, out_particle_f, dt)

    def coupling_prepare_proxy_contacts(self, state, contacts, *, contacts_freshly_detected: bool = False):
        return contacts  # no Contacts rows are read: boundaries come from shape SDFs and particle walls

    def coupling_supports_full_surface_soft_contacts(self) -> bool:
        return True

    def coupling_supports_inertial_property_refresh(self) -> bool:
        return False

    def alloy_restore_state_arrays(self, arrays: Mapping[str, np.ndarray]) -> None:
        for key, which in (("sph.kappa_v", _abi.BUF_KV), ("sph.kappa_d", _abi.BUF_KD)):
            self._write(which, np.ascontiguousarray(arrays[key], dtype=np.float32))
        if "sph.meter_j" in arrays:
            self._write(_abi.BUF_SHAPE_METER, np.ascontiguousarray(arrays["sph.meter_j"], dtype=np.int64))
        if "sph.phase" in arrays:
            self.model.sph.phase.copy_(torch.tensor(arrays["sph.phase"], device=self.device))
        if "sph.flags" in arrays:
            self.model.particle_flags.copy_(torch.tensor(arrays["sph.flags"], device=self.device))
        for key, which, dt in self._ledger_keys():
            if key in arrays:
                self._write(which, np.ascontiguousarray(arrays[key], dtype=dt))

    def alloy_state_worlds(self) -> dict[str, np.ndarray]:
        """Item -> world maps of the state arrays (coupling_hooks.py:128-138)."""
        w = self.model.particle_world.cpu().numpy()
        out = {"sph.kappa_v": w, "sph.kappa_d": w, "sph.phase": w, "sph.flags": w}
        if self.model.shape_count:
            out["sph.meter_j"] = np.asarray(self.model.shape_world, np.int32)
        wl = np.arange(self.model.world_count, dtype=np.int32)
        ew = np.asarray(self.emitters.table.world, np.int32)
        for key, _, _ in self._ledger_keys():
            out[key] = ew if key.startswith("sph.emit") else wl
        return out

    def export_flux_energy(self, health) -> None:
        """Energy emitted minus energy retired since the last call, into ``health``'s W_ext (solver.py:413-418,
        runtime.py:29-37 and 360-369): W_ext[w] = flux_e + flux_g z_ref[w], then the flux ledger is zeroed."""
        W = self.model.world_count
        fe = self.buffer(_abi.BUF_FLUX_E, np.int64, (W,))
        fg = self.buffer(_abi.BUF_FLUX_G, np.int64, (W,))
        z = health.state_arrays()["health.energy_z_ref"]
        z = np.asarray(z.detach().cpu().numpy() if isinstance(z, torch.Tensor) else
                       (z.numpy() if hasattr(z, "numpy") else z), np.float64)[:W]
        w_ext = fe.astype(np.float64) / _abi.WORK_SCALE + fg.astype(np.float64) / _abi.FORCE_SCALE * z
        for which in (_abi.BUF_FLUX_E, _abi.BUF_FLUX_G):
            self._write(which, np.zeros(W, np.int64))
        health.add_external_work(w_ext)

    def notify_model_changed(self, flags: int) -> None:
        """solver.py:343-351. Shape properties (transform, scale, radius, friction): the native handle is rebuilt from
        the model's shapes, keeping the warm starts, shape meters, tuning options and learned warm start. Model
        properties (gravity): the per-world gravity is re-read. This solver owns no bodies."""
        if flags & int(ModelFlags.MODEL_PROPERTIES):
            self._gravity.copy_(self.model.gravity.reshape(-1))
        if flags & int(ModelFlags.SHAPE_PROPERTIES):
            saved = self.alloy_state_arrays(None)
            W = self.model.world_count
            ledger = {which: self.buffer(which, dt, (W,)) for which, dt in (
                (_abi.BUF_MASS_INITIAL, np.int64), (_abi.BUF_EMIT_OVERFLOW, np.int32), (_abi.BUF_FLUX_E, np.int64),
                (_abi.BUF_FLUX_G, np.int64))}
            E = len(self.emitters.table.world)
            enabled = self.buffer(_abi.BUF_EMIT_ENABLED, np.int32, (E,)) if E else None
            S = self.model.shape_count
            meters = self.buffer(_abi.BUF_SHAPE_METER, np.int64, (max(S, 1), 3)) if S else None
            opts = {o: self.get_option(o) for o in _abi.ALL_OPTIONS}
            self.lib.dfsph_destroy(self._h)
            self._h = None
            self._create_native()
            if self.emitters.active:
                self.emitters.upload()
                if enabled is not None:
                    self._write(_abi.BUF_EMIT_ENABLED, enabled)
            for which, a in ledger.items():
                self._write(which, a)
            self.alloy_restore_state_arrays(saved)
            if meters is not None and self.model.shape_count == S:
                self._call(self.lib.dfsph_write_buffer, _abi.BUF_SHAPE_METER, meters.ctypes.data_as(ctypes.c_void_p),
                           meters.nbytes)
            if self._warm_model is not None:
                self._call(self.lib.dfsph_set_warm_start_model, _dptr(self._warm_model), None)
            for o, v in opts.items():
                try:
                    self.set_option(o, v)
                except ValueError:
                    if o != _abi.OPT_PACKED_ACCEL:  # the one option a new handle may refuse (several masses)
                        raise

    def alloy_shape_work(self, out: torch.Tensor) -> None:
        """Work [J] each shape's boundary did on the liquid in the last step at the wall's point velocity, per model
        shape (solver.py:407-412). The shapes this solver accepts are static, so the work is 0."""
        out.zero_()

    def retire(self, state: State, mask: torch.Tensor) -> None:
        """Retire the masked liquid particles for good (solver.py:428-431, ports.py:66-93): ACTIVE cleared, phase
        retired, velocity zeroed; their mass and energy leave the ledgers as at a sink."""
        n = self.model.particle_count
        if n == 0:
            return
        m32 = mask.to(device=self.device, dtype=torch.int32).contiguous()
        if tuple(m32.shape) != (n,):
            raise ValueError(f"SolverDFSPH.retire: mask must have {n} entries")
        state = device_state(state, self.device, n)
        self._checked_state(state, n)
        self._call(self.lib.dfsph_retire, _dptr(m32), _dptr(state.particle_q), _dptr(state.particle_qd),
                   _stream_ptr(self.device), what="SolverDFSPH.retire")

    def add_emitter(self, world: int, region, half_extents, rate: float, velocity, pool=None) -> int:
        """Emit liquid from a box region at ``rate`` [m^3/s] with ``velocity`` [m/s] by activating spares
        (solver.py:433-436, runtime.py:243-270); ``region`` is p(3) + q(4, xyzw). Returns the emitter's index."""
        return self.emitters.add_emitter(world, region, half_extents, rate, velocity, pool=pool)

    def set_emitter_enabled(self, e: int, on: bool) -> None:
        """Open or close emitter ``e`` between steps: a closed emitter emits nothing and its clock pauses
        (solver.py:438-447)."""
        n = len(self.emitters.table.world)
        if not 0 <= int(e) < n:
            raise ValueError(f"set_emitter_enabled: emitter {e} does not exist ({n} emitters)")
        self.emitters.flush()
        self._call(self.lib.dfsph_set_emitter_enabled, int(e), int(bool(on)))

    def add_sink(self, world: int, region, half_extents) -> int:
        """Retire particles that enter a box region; their mass goes to ``mass_out`` (solver.py:460-462)."""
        return self.emitters.add_sink(world, region, half_extents)

    def _no_patch(self, what: str):
        raise NotImplementedError(f"SolverDFSPH.{what}: patch emitters move with bodies (patch.py); bodies attached "
                                  "to shapes are not built in this solver")

    def add_patch_emitter(self, world_mask, body, local_points, velocity=(0.0, 0.0, 0.0)) -> int:
        self._no_patch("add_patch_emitter")

    def trigger(self, pid: int, world_flags) -> None:
        self._no_patch("trigger")

    @staticmethod
    def emit(builder, obj, decision) -> None:
        """Scene objects become liquid blocks (solver.py:475-479)."""
        from .build import add_fluid_block

        add_fluid_block(builder, decision.config, obj.lo, obj.counts)
