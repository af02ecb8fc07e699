"""ctypes binding of the DFSPH C-ABI (include/dfsph_abi.h) and the host-side tables it is created from.

``params_from_config`` derives, in float64 on the host, every construction-time constant the reference computes in
SolverDFSPH.__init__ (solver.py:92-94), shape_table (layout.py:195-249) and the tolerance driver
(solve_driver.py:88-91); the library rounds each to float32 where a Warp kernel argument would be float32.
``layout_from_model`` builds the world ranges, the per-world neighbour-grid boxes and the per-world shape table
(boundary.py:79-95).
"""

from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

from . import kernels as kn
from . import csf
from . import wall_viscosity as wv
from .config import DFSPHConfig, full_neighbor_count
from .model import GeoType, Model, ParticleFlags

ABI_VERSION = 5

DFSPH_OK, DFSPH_EINVAL, DFSPH_ECAPACITY, DFSPH_EDEVICE = 0, -1, -2, -3

(BUF_NBR, BUF_NBR_HI, BUF_COUNT, BUF_GROUP, BUF_SLOT_SHAPE, BUF_SLOT_PSI, BUF_SLOT_GPSI, BUF_SLOT_XB, BUF_SLOT_VB,
 BUF_SLOT_PHI, BUF_SLOT_XBL, BUF_SLOT_NBL, BUF_SLOT_IMPULSE, BUF_RATIO, BUF_ALPHA, BUF_NEFF, BUF_ACTIVE, BUF_REG,
 BUF_SRC, BUF_KAPPA, BUF_ACC, BUF_ERR, BUF_LOWER, BUF_KV, BUF_KD, BUF_VTMP, BUF_CONVERGED, BUF_ITERS, BUF_ERR_MEAN,
 BUF_ERR_MAX, BUF_N_ACTIVE, BUF_TOL_SCALE, BUF_OVERFLOW, BUF_MAX_SPEED, BUF_ITERS_DEN, BUF_ITERS_DIV, BUF_DEN_ERR,
 BUF_DIV_ERR, BUF_SHAPE_METER, BUF_FEATURES, BUF_MASS_INITIAL, BUF_MASS_IN, BUF_MASS_OUT, BUF_MASS_ACTIVE, BUF_ESCAPED,
 BUF_EMIT_OVERFLOW, BUF_FLUX_E, BUF_FLUX_G, BUF_EMIT_T, BUF_EMIT_CURSOR, BUF_EMIT_ENABLED, BUF_FLAGS,
 BUF_PHASE, BUF_PW_EXCH, BUF_PW_WORK, BUF_SKIN_FLAG, BUF_VISC_ITERS, BUF_CSF_SURF, BUF_CSF_KAPPA,
 BUF_PW_ROW_COUNT, BUF_HALO_COUNT) = range(61)
IMP_SCALE = float(2**50)    # impulse fixed point, bodies.py:25

FIX_SCALE = float(2**40)    # mass fixed point, kernels.py:32-33
WORK_SCALE = float(2**50)   # energy fixed point, bodies.py:24
FORCE_SCALE = float(2**40)  # bodies.py:23

# kernel ids of dfsph_profile_read
# tuning options (dfsph_set_option); none but OPT_LEARNED_SOLVES and OPT_LEARNED_GATE (which start each solve takes)
# changes a result bit
(OPT_CHECK_INTERVAL, OPT_GRAPHS, OPT_PAIR_CACHE, OPT_LEARNED_SOLVES, OPT_SPARSE_CORRECTION, OPT_BUILD_BLOCK,
 OPT_PACKED_ACCEL, OPT_HALO, OPT_DEVICE_LOOP, OPT_BUILD_STAGED, OPT_LEARNED_GATE) = (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11)
ALL_OPTIONS = (OPT_CHECK_INTERVAL, OPT_GRAPHS, OPT_PAIR_CACHE, OPT_LEARNED_SOLVES, OPT_SPARSE_CORRECTION,
               OPT_BUILD_BLOCK, OPT_PACKED_ACCEL, OPT_HALO, OPT_DEVICE_LOOP, OPT_BUILD_STAGED, OPT_LEARNED_GATE)
PAIR_CACHE_ACCEL, PAIR_CACHE_APPLY, PAIR_CACHE_OTHER, PAIR_CACHE_ALL = 1, 2, 4, 7

(KID_ACCEL, KID_APPLY, KID_BUILD, KID_DENSITY, KID_SLOTS, KID_NONPRESSURE, KID_ADVECT, KID_SOLVE_INIT,
 KID_POST) = range(9)
KERNEL_NAMES = ("k_accel", "k_apply", "k_build_lists", "k_density_alpha", "k_slots", "k_nonpressure",
                "k_advect_project", "k_solve_init", "k_post_solve")

PROJECT_EPS = 1.0e-3  # wall projection skin in particle radii (solver.py:483)
SWEEP = 0.5           # integrate.py:38
THIN_SWEEP = 0.5      # integrate.py:39
GRID_CELL_SLACK = 1.0e-5  # cell side h (1 + slack): pairs within h are always in adjacent cells
GRID_MAX_DIM = 1024


class CapacityError(RuntimeError):
    """A size beyond the solver's fixed capacity (alloy.core.errors.CapacityError)."""


c_int, c_double, c_float = ctypes.c_int32, ctypes.c_double, ctypes.c_float
P = ctypes.POINTER


class DfsphParams(ctypes.Structure):
    _fields_ = [
        ("particle_radius", c_double), ("h", c_double), ("spacing", c_double), ("kernel", c_int),
        ("pressure_solver", c_int), ("max_density_error", c_double), ("max_divergence_error", c_double),
        ("min_density_error", c_double), ("max_error_factor", c_double), ("rest_jitter_speed", c_double),
        ("rest_jitter_flow_ratio", c_double), ("min_iterations", c_int), ("max_iterations", c_int),
        ("min_neighbors_divergence", c_double), ("divergence_solve", c_int), ("divergence_two_sided", c_int),
        ("jacobi_omega", c_double), ("warm_start_density", c_double), ("warm_start_divergence", c_double),
        ("xsph_c", c_double), ("neighbor_capacity", c_int), ("wall_friction", c_int), ("n_full", c_double),
        ("tot_min", c_double), ("lattice_grad", c_double), ("compression_relief_speed", c_double),
        ("contact_separation_speed", c_double), ("wall_offset", c_double), ("wall_amplitude", c_double),
        ("taper0", c_double), ("taper1", c_double), ("reach", c_double), ("curvature", c_int),
        ("project_eps", c_double), ("sweep_len", c_double), ("one_way", c_int), ("skin", c_double),
        ("skin_rows", c_int), ("viscosity_implicit", c_int), ("visc_d", c_double), ("visc_tol", c_double),
        ("visc_max_iterations", c_int), ("boundary_viscosity", c_double), ("visc_wall_coeff", c_double),
        ("csf", c_int), ("csf_c_surface", c_double), ("csf_inv_i", c_double), ("csf_band_lo", c_double),
        ("csf_band_hi", c_double), ("csf_k_cohesion", c_double),
    ]


class DfsphLayout(ctypes.Structure):
    _fields_ = [
        ("n", c_int), ("world_count", c_int), ("world_start", P(c_int)), ("grid_origin", P(c_float)),
        ("grid_dims", P(c_int)), ("shape_count", c_int), ("shape_start", P(c_int)), ("shape_ids", P(c_int)),
        ("shape_geo", P(c_int)), ("shape_scale", P(c_float)), ("shape_xform", P(c_float)),
        ("shape_body", P(c_int)), ("shape_radius", P(c_float)), ("shape_mu", P(c_float)),
        ("wall_particle_count", c_int), ("shape_cos", P(c_float)), ("shape_sin", P(c_float)),
    ]


class DfsphModel(ctypes.Structure):
    _fields_ = [("mass", ctypes.c_void_p), ("rest_density", ctypes.c_void_p), ("flags", ctypes.c_void_p),
                ("phase", ctypes.c_void_p), ("world", ctypes.c_void_p), ("gravity", ctypes.c_void_p),
                ("wall_volume", ctypes.c_void_p), ("viscosity", ctypes.c_void_p),
                ("surface_tension", ctypes.c_void_p), ("csf_gamma", ctypes.c_void_p)]


class DfsphReport(ctypes.Structure):
    _fields_ = [("density_err", P(c_float)), ("divergence_err", P(c_float)), ("iters_density", P(c_int)),
                ("iters_divergence", P(c_int)), ("neighbor_overflow", P(c_int)), ("max_speed", P(c_float)),
                ("density_err_max", P(c_float)), ("density_tol", P(c_float)), ("projected", P(c_int)),
                ("swept", P(c_int)), ("nonfinite", P(c_int)), ("n_active", P(c_double))]


class DfsphEmitters(ctypes.Structure):
    _fields_ = [("emitter_count", c_int), ("world", P(c_int)), ("xform", P(c_float)), ("velocity", P(c_float)),
                ("interval", P(c_float)), ("offset_start", P(c_int)), ("offsets", P(c_float)),
                ("pool_start", P(c_int)), ("pool_ids", P(c_int)), ("sink_count", c_int), ("sink_world", P(c_int)),
                ("sink_xform", P(c_float)), ("sink_half", P(c_float)), ("has_bounds", c_int),
                ("bounds_lo", c_float * 3), ("bounds_hi", c_float * 3)]


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(P(ctype))


# ------------------------------------------------------------------------------------------------ parameters
def thinnest_wall(model: Model, ids: np.ndarray) -> float:
    """The thinnest primitive of the shape table (layout.py:128-148): 2 x the smallest box half extent, a sphere's
    diameter; planes count as thick."""
    out = math.inf
    for s in np.unique(ids):
        g = int(model.shape_type[s])
        if g == GeoType.BOX:
            out = min(out, 2.0 * float(np.min(model.shape_scale[s])))
        elif g == GeoType.SPHERE:
            out = min(out, 2.0 * float(model.shape_scale[s][0]))
    return out


def sweep_length(h: float, thinnest: float) -> float:
    """min(SWEEP h, THIN_SWEEP x thinnest wall) (integrate.py:42-47)."""
    return min(SWEEP * float(h), THIN_SWEEP * float(thinnest))


def params_from_config(cfg: DFSPHConfig, thinnest: float = math.inf) -> DfsphParams:
    sp, h, kind = cfg.spacing, cfg.h, cfg.kernel
    delta, amp = kn.boundary_fit(sp, h, kind)
    diag = kn.lattice_diagonal(sp, h, kind)
    reach = float(h) if cfg.boundary_viscosity else 1.5 * sp
    p = DfsphParams()
    p.particle_radius, p.h, p.spacing, p.kernel = cfg.particle_radius, h, sp, cfg.kernel_id
    p.pressure_solver = 0 if cfg.pressure_solver == "cg" else 1
    p.max_density_error, p.max_divergence_error = cfg.max_density_error, cfg.max_divergence_error
    p.min_density_error, p.max_error_factor = cfg.min_density_error, cfg.max_error_factor
    p.rest_jitter_speed = float(cfg.rest_jitter_speed) if cfg.rest_jitter_speed else 0.0
    p.rest_jitter_flow_ratio = cfg.rest_jitter_flow_ratio
    p.min_iterations, p.max_iterations = cfg.min_iterations, cfg.max_iterations
    p.min_neighbors_divergence = cfg.min_neighbors_divergence
    p.divergence_solve, p.divergence_two_sided = int(cfg.divergence_solve), int(cfg.divergence_two_sided)
    p.jacobi_omega = cfg.jacobi_omega
    p.warm_start_density, p.warm_start_divergence = cfg.warm_start_density, cfg.warm_start_divergence
    p.xsph_c = cfg.xsph_c
    p.neighbor_capacity, p.wall_friction = cfg.neighbor_capacity, int(bool(cfg.wall_friction))
    p.n_full = full_neighbor_count(cfg)
    p.tot_min = cfg.diagonal_floor * diag
    p.lattice_grad = math.sqrt(diag)
    p.compression_relief_speed = float(cfg.compression_relief_speed) if cfg.compression_relief_speed else 0.0
    p.contact_separation_speed = cfg.contact_separation_speed
    p.wall_offset, p.wall_amplitude = float(delta), float(amp)
    p.taper0, p.taper1, p.reach = 0.5 * sp, 1.5 * sp, reach
    p.curvature = int(cfg.boundary_curvature_order)
    p.project_eps = PROJECT_EPS * cfg.particle_radius
    p.sweep_len = sweep_length(h, thinnest)
    p.one_way = int(cfg.coupling == "one_way")
    p.skin = (cfg.reneighbor_skin or 0.0) * h
    p.skin_rows = skin_capacity(h, sp, p.skin) if p.skin > 0.0 else 0
    p.viscosity_implicit = 0  # set by the solver when the viscosity resolves to weiler2018
    p.visc_d = kn.LAPLACE_D / kn.laplacian_lattice_factor(sp, h, kind)
    p.visc_tol, p.visc_max_iterations = cfg.viscosity_tolerance, cfg.viscosity_max_iterations
    p.boundary_viscosity = float(cfg.boundary_viscosity or 0.0)
    p.visc_wall_coeff = (p.boundary_viscosity * wv.wall_coefficient_scale(sp, h, kind)
                         if p.boundary_viscosity > 0.0 else 0.0)
    p.csf = int(cfg.surface_tension == "csf")
    if p.csf:
        big_i, c_s = csf.flat_lattice_constants(sp, h, kind)
        p.csf_c_surface, p.csf_inv_i = c_s, 1.0 / big_i
        p.csf_band_lo, p.csf_band_hi = csf.CSF_BAND
        p.csf_k_cohesion = 32.0 / (math.pi * h**9)
    return p


SKIN_FILL_MARGIN = 1.75  # superset capacity / interior lattice neighbours within h + skin (neighbors_skin.py:146)


def skin_capacity(h: float, spacing: float, skin: float) -> int:
    """Superset row capacity (neighbors_skin.py:149-158): 1.75 x the neighbours of an interior lattice particle
    within h + skin, at least 16; this library holds at most 64 (a fuller row is built plain that step and its
    world rebuilds next step: reuse is lost, never exactness)."""
    rad = (h + skin) / spacing
    m = int(math.ceil(rad)) + 1
    ax = np.arange(-m, m + 1, dtype=np.float64)
    d2 = ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2
    fill = int(np.count_nonzero(d2 < rad * rad)) - 1
    return min(64, max(16, int(math.ceil(SKIN_FILL_MARGIN * fill))))


# ------------------------------------------------------------------------------------------------ layout
@dataclass
class LayoutTables:
    """Host arrays behind a DfsphLayout (kept alive while the struct is in use)."""

    world_start: np.ndarray
    grid_origin: np.ndarray
    grid_dims: np.ndarray
    shape_start: np.ndarray
    shape_ids: np.ndarray
    shape_geo: np.ndarray
    shape_scale: np.ndarray
    shape_xform: np.ndarray
    shape_body: np.ndarray
    shape_radius: np.ndarray
    shape_mu: np.ndarray
    wall_volume: np.ndarray
    wall_count: int
    thinnest: float
    shape_count: int = 0  # the model's shapes (the arrays above hold one placeholder entry when there are none)
    shape_cos: np.ndarray | None = None  # contact angles (csf), one placeholder entry when there are no shapes
    shape_sin: np.ndarray | None = None

    def struct(self, n: int) -> DfsphLayout:
        L = DfsphLayout()
        L.n, L.world_count = n, len(self.world_start) - 1
        L.world_start = _ptr(self.world_start, c_int)
        L.grid_origin, L.grid_dims = _ptr(self.grid_origin, c_float), _ptr(self.grid_dims, c_int)
        L.shape_count = int(self.shape_count)
        L.shape_start, L.shape_ids = _ptr(self.shape_start, c_int), _ptr(self.shape_ids, c_int)
        L.shape_geo, L.shape_scale = _ptr(self.shape_geo, c_int), _ptr(self.shape_scale, c_float)
        L.shape_xform, L.shape_body = _ptr(self.shape_xform, c_float), _ptr(self.shape_body, c_int)
        L.shape_radius, L.shape_mu = _ptr(self.shape_radius, c_float), _ptr(self.shape_mu, c_float)
        L.wall_particle_count = int(self.wall_count)
        if self.shape_cos is not None:
            L.shape_cos, L.shape_sin = _ptr(self.shape_cos, c_float), _ptr(self.shape_sin, c_float)
        return L


def world_shape_table(shape_world: np.ndarray, shape_flags: np.ndarray, shape_type: np.ndarray,
                      world_count: int) -> tuple[np.ndarray, np.ndarray]:
    """CSR (start[W+1], ids) of particle-colliding shapes per world, global shapes appended (boundary.py:79-95)."""
    from .model import ShapeFlags

    usable = ((shape_flags & int(ShapeFlags.COLLIDE_PARTICLES)) != 0) & np.isin(
        shape_type, [int(GeoType.PLANE), int(GeoType.SPHERE), int(GeoType.BOX)])
    glob = np.nonzero(usable & (shape_world < 0))[0]
    start, ids = [0], []
    for w in range(world_count):
        ids.extend(int(s) for s in np.nonzero(usable & (shape_world == w))[0])
        ids.extend(int(s) for s in glob)
        start.append(len(ids))
    return np.asarray(start, np.int32), np.asarray(ids, np.int32)


def grid_boxes(q: np.ndarray, world_start: np.ndarray, model: Model, ids_per_world: list[np.ndarray], h: float,
               bounds=None) -> tuple[np.ndarray, np.ndarray]:
    """Per-world cell-grid origin and dims covering the world's particles (spares too) and its boxes/spheres, padded
    by 2 h, and ``bounds`` when given. The grid only accelerates the search: positions outside it are clamped to its
    edge cells, which keeps every pair within h in adjacent cells; the escape box is left out by the caller (an
    automatic one is padded by at least 0.25 m and would multiply the cells)."""
    W = len(world_start) - 1
    cell = h * (1.0 + GRID_CELL_SLACK)
    org = np.zeros((W, 3), np.float32)
    dims = np.ones((W, 3), np.int32)
    for w in range(W):
        lo = np.full(3, np.inf)
        hi = np.full(3, -np.inf)
        pts = q[world_start[w]:world_start[w + 1]]
        pts = pts[np.all(np.isfinite(pts), axis=1)]
        if len(pts):
            lo, hi = np.minimum(lo, pts.min(0)), np.maximum(hi, pts.max(0))
        for s in ids_per_world[w]:
            g = int(model.shape_type[s])
            if g == GeoType.PLANE:
                continue
            c = model.shape_transform[s, :3].astype(np.float64)
            r = float(model.shape_collision_radius[s])
            if g == GeoType.BOX and np.allclose(model.shape_transform[s, 3:], [0, 0, 0, 1]):
                e = model.shape_scale[s].astype(np.float64)
            else:
                e = np.full(3, r)
            lo, hi = np.minimum(lo, c - e), np.maximum(hi, c + e)
        if bounds is not None:
            lo, hi = np.minimum(lo, np.asarray(bounds[0])), np.maximum(hi, np.asarray(bounds[1]))
        if not np.all(np.isfinite(lo)):
            lo, hi = np.zeros(3), np.zeros(3)
        lo, hi = lo - 2.0 * h, hi + 2.0 * h
        org[w] = lo.astype(np.float32)
        dims[w] = np.clip(np.ceil((hi - lo) / cell).astype(np.int64) + 1, 1, GRID_MAX_DIM)
    return org, dims


def layout_from_model(model: Model, cfg: DFSPHConfig) -> LayoutTables:
    W = model.world_count
    ws = np.asarray(model.world_start_np, np.int32)
    start, ids = world_shape_table(model.shape_world, model.shape_flags, model.shape_type, W) if model.shape_count \
        else (np.zeros(W + 1, np.int32), np.zeros(0, np.int32))
    flags = model.particle_flags.cpu().numpy()
    radius = model.particle_radius.cpu().numpy().astype(np.float64)
    phase = model.sph.phase.cpu().numpy()
    wall = ((flags & int(ParticleFlags.PROXY)) != 0) & (phase < 0)
    vol = np.zeros(model.particle_count, np.float32)
    vol[wall] = (8.0 * radius[wall] ** 3).astype(np.float32)
    q = model.particle_q.cpu().numpy().astype(np.float64)
    per_world = [ids[start[w]:start[w + 1]] for w in range(W)]
    org, dims = grid_boxes(q, ws, model, per_world, cfg.h + (cfg.reneighbor_skin or 0.0) * cfg.h)
    S = model.shape_count
    cos, sin = csf.shape_angles(model.sph.shape_contact_angle if S else np.zeros(0))
    f32 = lambda a, shape: np.ascontiguousarray(np.asarray(a, np.float32).reshape(shape))  # noqa: E731
    i32 = lambda a: np.ascontiguousarray(np.asarray(a, np.int32).reshape(-1))  # noqa: E731
    return LayoutTables(
        world_start=i32(ws), grid_origin=f32(org, -1), grid_dims=i32(dims), shape_start=i32(start),
        shape_ids=i32(ids if len(ids) else np.zeros(1)), shape_geo=i32(model.shape_type if S else np.zeros(1)),
        shape_scale=f32(model.shape_scale if S else np.zeros(3), -1),
        shape_xform=f32(model.shape_transform if S else np.zeros(7), -1),
        shape_body=i32(model.shape_body if S else np.full(1, -1)),
        shape_radius=f32(model.shape_collision_radius if S else np.zeros(1), -1),
        shape_mu=f32(model.shape_material_mu if S else np.zeros(1), -1),
        wall_volume=vol, wall_count=int(np.count_nonzero(wall)),
        thinnest=thinnest_wall(model, ids) if S else math.inf, shape_count=S,
        shape_cos=f32(cos, -1), shape_sin=f32(sin, -1),
    )


# ------------------------------------------------------------------------------------------------ library
_LIB_NAME = "libdfsph_hip.so"


def library_path() -> str:
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "csrc", _LIB_NAME)


_lib = None


def load_library() -> ctypes.CDLL:
    """The HIP library, loaded once. There is no fallback: a missing library is an error."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"DFSPH HIP library not built: {path} (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "dfsph_abi_version": ([], i32), "dfsph_last_error": ([], ctypes.c_char_p),
        "dfsph_create": ([P(DfsphParams), P(DfsphLayout), P(DfsphModel), P(vp)], i32),
        "dfsph_destroy": ([vp], i32),
        "dfsph_step": ([vp, vp, vp, vp, dbl, vp], i32),
        "dfsph_build_neighbors": ([vp, vp, vp], i32),
        "dfsph_boundary_slots": ([vp, vp, vp, vp], i32),
        "dfsph_density_alpha": ([vp, vp, vp], i32),
        "dfsph_flow_speed": ([vp, vp, vp], i32),
        "dfsph_solve": ([vp, i32, vp, vp, dbl, i32, vp], i32),
        "dfsph_nonpressure": ([vp, vp, vp, vp, dbl, vp], i32),
        "dfsph_advect_project": ([vp, vp, vp, dbl, vp], i32),
        "dfsph_prime": ([vp, vp, vp, vp, dbl, dbl, i32, P(i32), vp], i32),
        "dfsph_set_warm_start_model": ([vp, vp, vp], i32),
        "dfsph_set_warm_start_kappa": ([vp, vp, vp], i32),
        "dfsph_set_check_interval": ([vp, i32], i32),
        "dfsph_set_graphs": ([vp, i32], i32),
        "dfsph_set_option": ([vp, i32, i32], i32),
        "dfsph_get_option": ([vp, i32, P(i32)], i32),
        "dfsph_mlp_param_count": ([], i32),
        "dfsph_mlp_forward": ([vp, vp, i32, vp, vp], i32),
        "dfsph_mlp_grad_scratch_floats": ([i32], i64),
        "dfsph_mlp_grad": ([vp, vp, vp, i32, i32, vp, i64, vp, vp], i32),
        "dfsph_mlp_grad_noalloc": ([vp, vp, vp, i32, i32, vp, vp], i32),
        "dfsph_mlp_adam": ([vp, vp, vp, vp, dbl, dbl, dbl, dbl, i32, vp], i32),
        "dfsph_set_logging": ([vp, vp, vp], i32),
        "dfsph_get_buffer": ([vp, i32, P(vp), P(i64)], i32),
        "dfsph_set_profiling": ([vp, i32, i32], i32),
        "dfsph_profile_read": ([vp, i32, P(i32), P(dbl), P(dbl)], i32),
        "dfsph_read_buffer": ([vp, i32, vp, i64], i32),
        "dfsph_write_buffer": ([vp, i32, vp, i64], i32),
        "dfsph_zero_stats": ([vp, vp], i32),
        "dfsph_report": ([vp, P(DfsphReport), vp], i32),
        "dfsph_set_emitters": ([vp, P(DfsphEmitters)], i32),
        "dfsph_set_emitter_enabled": ([vp, i32, i32], i32),
        "dfsph_emitters_step": ([vp, vp, vp, dbl, vp], i32),
        "dfsph_active_mass": ([vp, vp], i32),
        "dfsph_retire": ([vp, vp, vp, vp, vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes, fn.restype = args, res
    if lib.dfsph_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{path}: ABI version {lib.dfsph_abi_version()} != {ABI_VERSION}")
    _lib = lib
    return lib


def check(rc: int, lib, what: str = "dfsph") -> None:
    """Raise the Python error a status maps to (ValueError, CapacityError, RuntimeError)."""
    if rc == DFSPH_OK:
        return
    msg = f"{what}: {lib.dfsph_last_error().decode()}"
    if rc == DFSPH_EINVAL:
        raise ValueError(msg)
    if rc == DFSPH_ECAPACITY:
        raise CapacityError(msg)
    raise RuntimeError(msg)
