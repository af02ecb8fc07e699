"""Scene builders for SPH liquids: lattice blocks, spare pools, SDF box tanks and particle walls.

Reference: alloy/solvers/sph/build.py:26-137 (lattice_points, add_fluid_points, add_fluid_block, add_spares,
wall_box, add_box_tank) and particle_walls.py:201-240 (wall particles: ACTIVE|PROXY, sph:phase -1, volume 8 r^3).
Particle mass is rho0 / S (config.particle_mass), so an interior lattice particle sits exactly at rho0.
"""

from __future__ import annotations

import numpy as np

from .config import PHASE_LIQUID, PHASE_NONE, PHASE_SPARE, DFSPHConfig, particle_mass, register_custom_attributes
from .model import GeoType, ModelBuilder, ParticleFlags

Vec3 = tuple[float, float, float]


def lattice_points(lo: Vec3, counts: tuple[int, int, int], spacing: float) -> np.ndarray:
    """Cell centres (i + 1/2) spacing of an nx x ny x nz lattice from corner ``lo`` (build.py:26-31), x slowest."""
    axes = [lo[k] + (np.arange(c) + 0.5) * spacing for k, c in enumerate(counts)]
    gx, gy, gz = np.meshgrid(*axes, indexing="ij")
    return np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)


def add_fluid_points(builder: ModelBuilder, cfg: DFSPHConfig, points: np.ndarray, rest_density: float = 1000.0,
                     viscosity: float = 1.0e-3, velocity: Vec3 = (0.0, 0.0, 0.0), surface_tension: float = 0.0,
                     adhesion: float = 0.0) -> range:
    """Active liquid particles at ``points`` in the builder's current world (build.py:41-57); ``surface_tension``
    [N/m] acts with ``surface_tension="csf"``."""
    register_custom_attributes(builder)
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    if len(pts) == 0:
        return range(builder.particle_count, builder.particle_count)
    return builder.add_particles(pts, velocity, particle_mass(cfg, rest_density), cfg.particle_radius,
                                 int(ParticleFlags.ACTIVE), rest_density, viscosity, PHASE_LIQUID, surface_tension,
                                 adhesion)


def add_fluid_block(builder: ModelBuilder, cfg: DFSPHConfig, lo: Vec3, counts: tuple[int, int, int],
                    rest_density: float = 1000.0, viscosity: float = 1.0e-3, velocity: Vec3 = (0.0, 0.0, 0.0),
                    jitter: float = 0.0, rng: np.random.Generator | None = None) -> range:
    """A liquid block on the 2r lattice with its lower corner at ``lo`` (build.py:60-64), plus an optional uniform
    jitter in [-jitter, jitter] per coordinate from ``rng``."""
    pts = lattice_points(lo, counts, cfg.spacing)
    if jitter > 0.0:
        pts = pts + (rng or np.random.default_rng(0)).uniform(-jitter, jitter, pts.shape)
    return add_fluid_points(builder, cfg, pts, rest_density, viscosity, velocity)


def add_spares(builder: ModelBuilder, cfg: DFSPHConfig, count: int, rest_density: float = 1000.0,
               viscosity: float = 1.0e-3, park: Vec3 = (0.0, 0.0, 0.0)) -> range:
    """Inactive spare particles (ACTIVE cleared, phase 0) parked at ``park`` (build.py:67-81)."""
    register_custom_attributes(builder)
    if count <= 0:
        return range(builder.particle_count, builder.particle_count)
    pts = np.tile(np.asarray(park, np.float64), (count, 1))
    return builder.add_particles(pts, (0.0, 0.0, 0.0), particle_mass(cfg, rest_density), cfg.particle_radius, 0,
                                 rest_density, viscosity, PHASE_SPARE)


def wall_box(builder: ModelBuilder, center: Vec3, half: Vec3, mu: float = 0.5,
             contact_angle: float | None = None) -> int:
    """One static SDF wall box (build.py:91-110); ``contact_angle`` [rad] for csf (None: a free contact line)."""
    register_custom_attributes(builder)
    return builder.add_shape(GeoType.BOX, center, half, mu=mu,
                             contact_angle=-1.0 if contact_angle is None else float(contact_angle))


def add_box_tank(builder: ModelBuilder, inner_lo: Vec3, inner_hi: Vec3, wall: float, *, open_top: bool = True,
                 floor: bool = True, mu: float = 0.5) -> dict[str, int]:
    """Axis-aligned tank of box walls ``wall`` thick around the inner box (build.py:113-137): the floor spans the full
    footprint, x walls the full y extent, y walls sit between them, so the solid is tiled without overlap."""
    register_custom_attributes(builder)
    lo, hi = np.asarray(inner_lo, float), np.asarray(inner_hi, float)
    c = 0.5 * (lo + hi)
    hx, hy, hz = 0.5 * (hi - lo)
    t = 0.5 * wall
    ids = {}
    if floor:
        ids["floor"] = wall_box(builder, (c[0], c[1], lo[2] - t), (hx + wall, hy + wall, t), mu)
    for name, sx in (("x_lo", -1.0), ("x_hi", 1.0)):
        ids[name] = wall_box(builder, (c[0] + sx * (hx + t), c[1], c[2]), (t, hy + wall, hz), mu)
    for name, sy in (("y_lo", -1.0), ("y_hi", 1.0)):
        ids[name] = wall_box(builder, (c[0], c[1] + sy * (hy + t), c[2]), (hx, t, hz), mu)
    if not open_top:
        ids["lid"] = wall_box(builder, (c[0], c[1], hi[2] + t), (hx + wall, hy + wall, t), mu)
    return ids


def add_wall_particles(builder: ModelBuilder, cfg: DFSPHConfig, points: np.ndarray,
                       velocity: Vec3 = (0.0, 0.0, 0.0)) -> range:
    """Particle-wall particles (particle_walls.py:1-27, 201-209): ACTIVE|PROXY, sph:phase -1 (never liquid),
    radius r so their volume is 8 r^3 = spacing^3. They sit in the neighbour grid as group -2 - world and are
    read through the colour field only."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    return builder.add_particles(pts, velocity, particle_mass(cfg, 1000.0), cfg.particle_radius,
                                 int(ParticleFlags.ACTIVE | ParticleFlags.PROXY), 1000.0, 1.0e-3, PHASE_NONE)


def bowl_points(center: Vec3, radius: float, spacing: float, layers: int = 2) -> np.ndarray:
    """Lattice sites of a hemispherical shell (open top) of inner radius ``radius`` and ``layers`` lattice layers,
    centred at ``center`` (the rim at center.z): sites with radius <= |x - c| < radius + layers spacing and z <= c_z.
    Two layers make the shell at least h = 2 spacings thick, so its colour reaches the half-space value."""
    c = np.asarray(center, np.float64)
    outer = radius + layers * spacing
    m = int(np.ceil(outer / spacing)) + 1
    ax = (np.arange(-m, m + 1)) * spacing
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    d = np.linalg.norm(g, axis=1)
    keep = (d >= radius) & (d < outer) & (g[:, 2] <= 0.0)
    return g[keep] + c


def box_lining_points(inner_lo: Vec3, inner_hi: Vec3, spacing: float, layers: int = 2,
                      open_top: bool = True) -> np.ndarray:
    """Lattice sites lining a box (floor and four sides, ``layers`` thick, outside the inner box)."""
    lo, hi = np.asarray(inner_lo, np.float64), np.asarray(inner_hi, np.float64)
    t = layers * spacing
    counts = np.ceil((hi - lo + 2 * t) / spacing).astype(int)
    pts = lattice_points(tuple(lo - t), tuple(int(c) for c in counts), spacing)
    inside = np.all((pts > lo) & (pts < hi), axis=1)
    if open_top:
        inside |= np.all((pts[:, :2] > lo[:2]) & (pts[:, :2] < hi[:2]), axis=1) & (pts[:, 2] > lo[2])
        pts = pts[pts[:, 2] < hi[2]]
        inside = np.all((pts > lo) & (pts < hi), axis=1) | (
            np.all((pts[:, :2] > lo[:2]) & (pts[:, :2] < hi[:2]), axis=1) & (pts[:, 2] > lo[2]))
    return pts[~inside]
