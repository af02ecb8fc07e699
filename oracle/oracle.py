"""ctypes wrapper of the CPU oracle (liboracle_dfsph.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() (as the checker) and bench.py's cpu_baseline leg (as the timed CPU
baseline); never by the product (alloy_amd/). It builds the oracle from the same model, configuration, parameter
and layout tables the product's SolverDFSPH uses (the boundary structs of include/dfsph_abi.h), holds its own host
copy of the state, and exposes every solver buffer as a numpy view.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from alloy_amd.sph import _abi, csf
from alloy_amd.sph.config import DFSPHConfig
from alloy_amd.sph.model import Model

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_dfsph.so")
OMP_LIB_PATH = os.path.join(_HERE, "liboracle_dfsph_omp.so")  # the same oracle, loops over OpenMP threads
_libs: dict[bool, ctypes.CDLL] = {}


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ only)."""
    args = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True, capture_output=True)
    return LIB_PATH


def load(parallel: bool = False) -> ctypes.CDLL:
    """The serial oracle, or with ``parallel`` its OpenMP build (OMP_NUM_THREADS threads; the same results)."""
    if parallel in _libs:
        return _libs[parallel]
    path = OMP_LIB_PATH if parallel else LIB_PATH
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "orc_last_error": ([], ctypes.c_char_p),
        "orc_create": ([i32, P(_abi.DfsphParams), P(_abi.DfsphLayout), P(_abi.DfsphModel), P(vp)], i32),
        "orc_destroy": ([vp], i32),
        "orc_step": ([vp, vp, vp, vp, dbl, vp, vp], i32),
        "orc_prime": ([vp, vp, vp, vp, dbl, dbl, i32, P(i32)], i32),
        "orc_build_neighbors": ([vp, vp], i32),
        "orc_boundary_slots": ([vp, vp, vp], i32),
        "orc_density_alpha": ([vp, vp], i32),
        "orc_solve": ([vp, i32, vp, vp, dbl, i32], i32),
        "orc_flow_speed": ([vp, vp], i32),
        "orc_nonpressure": ([vp, vp, vp, vp, dbl], i32),
        "orc_advect_project": ([vp, vp, vp, dbl], i32),
        "orc_meters": ([vp], i32),
        "orc_zero_impulse": ([vp], i32),
        "orc_set_emitters": ([vp, P(_abi.DfsphEmitters)], i32),
        "orc_set_emitter_enabled": ([vp, i32, i32], i32),
        "orc_emitters_step": ([vp, vp, vp, dbl], i32),
        "orc_active_mass": ([vp], i32),
        "orc_get_buffer": ([vp, i32, P(vp), P(i64)], i32),
        "orc_reduce": ([vp, vp, i32, P(dbl), vp], i32),
        "orc_kernel_w": ([i32, dbl, dbl, i32], dbl),
        "orc_kernel_grad": ([i32, P(dbl), dbl, i32, P(dbl)], None),
        "orc_halfspace_volume": ([i32, dbl, dbl, dbl, i32, P(dbl)], None),
        "orc_shape_sdf": ([i32, P(dbl), P(dbl), P(dbl)], None),
        "orc_colour_distance": ([vp, dbl], dbl),
        "orc_shape_insert": ([vp, vp, vp, i32, i32, ctypes.c_float, vp, ctypes.c_float], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes, fn.restype = args, res
    _libs[parallel] = lib
    return lib


def _check(rc: int, what: str, lib=None) -> None:
    if rc != 0:
        msg = f"{what}: {(lib or load()).orc_last_error().decode()}"
        raise ValueError(msg) if rc == -1 else RuntimeError(msg)


_BUF_DTYPES = {
    _abi.BUF_NBR: np.uint16, _abi.BUF_NBR_HI: np.uint8, _abi.BUF_COUNT: np.int32, _abi.BUF_GROUP: np.int32,
    _abi.BUF_SLOT_SHAPE: np.int32, _abi.BUF_CONVERGED: np.int32, _abi.BUF_ITERS: np.int32,
    _abi.BUF_N_ACTIVE: np.float64, _abi.BUF_OVERFLOW: np.int32, _abi.BUF_ITERS_DEN: np.int32,
    _abi.BUF_ITERS_DIV: np.int32, _abi.BUF_SHAPE_METER: np.int64, _abi.BUF_MASS_INITIAL: np.int64,
    _abi.BUF_MASS_IN: np.int64, _abi.BUF_MASS_OUT: np.int64, _abi.BUF_MASS_ACTIVE: np.int64, _abi.BUF_ESCAPED: np.int32,
    _abi.BUF_EMIT_OVERFLOW: np.int32, _abi.BUF_FLUX_E: np.int64, _abi.BUF_FLUX_G: np.int64,
    _abi.BUF_EMIT_CURSOR: np.int32, _abi.BUF_EMIT_ENABLED: np.int32, _abi.BUF_FLAGS: np.int32, _abi.BUF_PHASE: np.int8,
    _abi.BUF_PW_EXCH: np.int64, _abi.BUF_PW_WORK: np.int64, _abi.BUF_VISC_ITERS: np.int32, _abi.BUF_CSF_SURF: np.int32,
}


class OracleDFSPH:
    """The CPU oracle for one model: ``precision`` 32 (bitwise parity with the HIP path) or 64 (FP64 reference)."""

    def __init__(self, model: Model, cfg: DFSPHConfig, precision: int = 32, pressure_solver: str | None = None,
                 viscosity: str | None = None, parallel: bool = False):
        self.lib = load(parallel)
        self.model, self.cfg, self.precision = model, cfg, precision
        self.real = np.float64 if precision == 64 else np.float32
        self.n, self.W = model.particle_count, model.world_count
        self.tables = _abi.layout_from_model(model, cfg)
        self.params = _abi.params_from_config(cfg, self.tables.thinnest)
        if pressure_solver is not None:
            self.params.pressure_solver = 0 if pressure_solver == "cg" else 1
        from alloy_amd.sph.config import resolve_viscosity

        visc = viscosity or resolve_viscosity(cfg, float(model.sph.viscosity.max().item()) if self.n else 0.0)
        self.params.xsph_c = csf.xsph_coefficient(cfg, visc)
        self.params.viscosity_implicit = int(visc == "weiler2018")
        host = lambda t, dt: np.ascontiguousarray(t.detach().cpu().numpy().astype(dt))  # noqa: E731
        self._keep = dict(
            mass=host(model.particle_mass, np.float32), rho0=host(model.sph.rest_density, np.float32),
            flags=host(model.particle_flags, np.int32), phase=host(model.sph.phase, np.int8),
            world=host(model.particle_world, np.int32), gravity=host(model.gravity.reshape(-1), np.float32),
            wvol=np.ascontiguousarray(self.tables.wall_volume, np.float32),
            mu=host(model.sph.viscosity, np.float32), sigma=host(model.sph.surface_tension, np.float32),
            gamma=np.ascontiguousarray(csf.liquid_gamma(model, cfg), np.float32))
        k = self._keep
        m = _abi.DfsphModel()
        for field_, key in (("mass", "mass"), ("rest_density", "rho0"), ("flags", "flags"), ("phase", "phase"),
                            ("world", "world"), ("gravity", "gravity"), ("wall_volume", "wvol"), ("viscosity", "mu"),
                            ("surface_tension", "sigma"), ("csf_gamma", "gamma")):
            setattr(m, field_, k[key].ctypes.data_as(ctypes.c_void_p))
        layout = self.tables.struct(self.n)
        h = ctypes.c_void_p()
        _check(self.lib.orc_create(precision, ctypes.byref(self.params), ctypes.byref(layout), ctypes.byref(m),
                                   ctypes.byref(h)), "orc_create")
        self.h = h
        self.q = np.ascontiguousarray(model.particle_q.detach().cpu().numpy().astype(self.real))
        self.qd = np.ascontiguousarray(model.particle_qd.detach().cpu().numpy().astype(self.real))

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.orc_destroy(self.h)
            self.h = None

    def _p(self, a):
        return None if a is None else a.ctypes.data_as(ctypes.c_void_p)

    def set_state(self, q: np.ndarray, qd: np.ndarray) -> None:
        self.q[...] = q
        self.qd[...] = qd

    def step(self, dt: float, f: np.ndarray | None = None, warm_div=None, warm_den=None) -> None:
        conv = lambda a: None if a is None else np.ascontiguousarray(a, self.real)  # noqa: E731
        f, warm_div, warm_den = conv(f), conv(warm_div), conv(warm_den)
        _check(self.lib.orc_step(self.h, self._p(self.q), self._p(self.qd), self._p(f), float(dt), self._p(warm_div),
                                 self._p(warm_den)), "orc_step")

    def prime(self, guess: np.ndarray, dt: float, tol: float = 1.0e-6, max_iterations: int = 2000) -> int:
        g = np.ascontiguousarray(guess, self.real)
        it = ctypes.c_int32(0)
        _check(self.lib.orc_prime(self.h, self._p(self.q), self._p(self.qd), self._p(g), float(dt), float(tol),
                                  int(max_iterations), ctypes.byref(it)), "orc_prime")
        return int(it.value)

    def build_neighbors(self):
        _check(self.lib.orc_build_neighbors(self.h, self._p(self.q)), "orc_build_neighbors")

    def boundary_slots(self):
        _check(self.lib.orc_boundary_slots(self.h, self._p(self.q), self._p(self.qd)), "orc_boundary_slots")

    def density_alpha(self):
        _check(self.lib.orc_density_alpha(self.h, self._p(self.q)), "orc_density_alpha")

    def flow_speed(self):
        _check(self.lib.orc_flow_speed(self.h, self._p(self.qd)), "orc_flow_speed")

    def solve(self, mode: int, dt: float, fixed_iters: int = -1):
        _check(self.lib.orc_solve(self.h, int(mode), self._p(self.q), self._p(self.qd), float(dt), int(fixed_iters)),
               "orc_solve")

    def nonpressure(self, dt: float, f: np.ndarray | None = None):
        f = None if f is None else np.ascontiguousarray(f, self.real)
        _check(self.lib.orc_nonpressure(self.h, self._p(self.q), self._p(self.qd), self._p(f), float(dt)),
               "orc_nonpressure")

    def meters(self):
        _check(self.lib.orc_meters(self.h), "orc_meters")

    def advect_project(self, dt: float):
        _check(self.lib.orc_advect_project(self.h, self._p(self.q), self._p(self.qd), float(dt)), "orc_advect_project")

    def set_emitters(self, table) -> None:
        """The emitter, sink and escape-box table (an ``alloy_amd.sph.emitters.EmitterStruct``, e.g. a solver's
        ``emitters.struct``)."""
        self._emit_keep = table
        _check(self.lib.orc_set_emitters(self.h, ctypes.byref(table.struct)), "orc_set_emitters")

    def set_emitter_enabled(self, e: int, on: bool) -> None:
        _check(self.lib.orc_set_emitter_enabled(self.h, int(e), int(bool(on))), "orc_set_emitter_enabled")

    def emitters_step(self, dt: float) -> None:
        _check(self.lib.orc_emitters_step(self.h, self._p(self.q), self._p(self.qd), float(dt)), "orc_emitters_step")

    def mass_ledger(self) -> dict[str, np.ndarray]:
        _check(self.lib.orc_active_mass(self.h), "orc_active_mass")
        return {"initial": self.buffer(_abi.BUF_MASS_INITIAL).copy(), "in": self.buffer(_abi.BUF_MASS_IN).copy(),
                "out": self.buffer(_abi.BUF_MASS_OUT).copy(), "active": self.buffer(_abi.BUF_MASS_ACTIVE).copy()}

    def buffer(self, which: int) -> np.ndarray:
        """A numpy view of an oracle buffer (float buffers in the oracle's precision)."""
        ptr, nb = ctypes.c_void_p(), ctypes.c_int64()
        _check(self.lib.orc_get_buffer(self.h, which, ctypes.byref(ptr), ctypes.byref(nb)), "orc_get_buffer")
        dt = np.dtype(_BUF_DTYPES.get(which, self.real))
        if which in (_abi.BUF_ERR_MEAN, _abi.BUF_ERR_MAX, _abi.BUF_TOL_SCALE, _abi.BUF_MAX_SPEED, _abi.BUF_DEN_ERR,
                     _abi.BUF_DIV_ERR, _abi.BUF_EMIT_T):
            dt = np.dtype(self.real)
        count = nb.value // dt.itemsize
        if count == 0:
            return np.zeros(0, dt)
        arr = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))), (count,))
        return arr

    def reduce(self, values: np.ndarray, ncomp: int) -> tuple[np.ndarray, np.ndarray]:
        v = np.ascontiguousarray(values, self.real).reshape(-1)
        out = np.zeros(self.W * ncomp, np.float64)
        mx = np.zeros(self.W, self.real)
        _check(self.lib.orc_reduce(self.h, self._p(v), ncomp, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   self._p(mx)), "orc_reduce")
        return out.reshape(self.W, ncomp), mx


def kernel_w(r: float, h: float, kind: int = 0, precision: int = 64) -> float:
    return load().orc_kernel_w(precision, r, h, kind)


def kernel_grad(rv, h: float, kind: int = 0, precision: int = 64) -> np.ndarray:
    a = np.ascontiguousarray(rv, np.float64)
    out = np.zeros(3)
    P = ctypes.POINTER(ctypes.c_double)
    load().orc_kernel_grad(precision, a.ctypes.data_as(P), h, kind, out.ctypes.data_as(P))
    return out


def halfspace_volume(d: float, h: float, H: float = 0.0, kind: int = 0, precision: int = 64) -> tuple[float, float]:
    out = np.zeros(2)
    load().orc_halfspace_volume(precision, d, h, H, kind, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return float(out[0]), float(out[1])


def shape_insert(slots, new, near: float) -> tuple[int, list[int], np.ndarray]:
    """Insert ``new`` = (shape, phi, normal) into the slot row ``slots`` (a list of up to 4 (shape, phi, normal),
    phi-ascending) with the oracle's float32 shape_insert; (kept, shapes after, phis after)."""
    shape = np.full(4, -1, np.int32)
    phi = np.zeros(4, np.float32)
    nrm = np.zeros((4, 3), np.float32)
    for k, (s_, p_, n_) in enumerate(slots):
        shape[k], phi[k], nrm[k] = s_, p_, n_
    nn = np.ascontiguousarray(new[2], np.float32)
    kept = load().orc_shape_insert(shape.ctypes.data_as(ctypes.c_void_p), phi.ctypes.data_as(ctypes.c_void_p),
                                   nrm.ctypes.data_as(ctypes.c_void_p), len(slots), int(new[0]), float(new[1]),
                                   nn.ctypes.data_as(ctypes.c_void_p), float(near))
    return int(kept), shape.tolist(), phi.copy()


def shape_sdf(geo: int, scale, p) -> tuple[float, np.ndarray]:
    P = ctypes.POINTER(ctypes.c_double)
    s = np.ascontiguousarray(scale, np.float64)
    x = np.ascontiguousarray(p, np.float64)
    out = np.zeros(4)
    load().orc_shape_sdf(geo, s.ctypes.data_as(P), x.ctypes.data_as(P), out.ctypes.data_as(P))
    return float(out[0]), out[1:].copy()
