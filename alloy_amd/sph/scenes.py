"""Seeded reference scenes: resting columns, dam breaks and batched bowl worlds (alloy/testing/sph.py:49-115).

The benchmark workloads are named by what they are:
- ``dam_break(n_side, box, seed)``: an n_side^3 lattice block (spacing 0.01 m, jitter U(+-1e-4) m from ``seed``) in
  one corner of a cubic SDF box tank of side ``box`` with walls h thick. 20^3 in 0.4 m and 50^3 in 1.0 m are the
  8,000- and 125,000-particle dam breaks; 100^3 in 2.0 m the 1,000,000-particle one.
- ``bowl_worlds(worlds, seed)``: batched worlds, each a box tank of side U(0.3, 0.6) m with a static particle-wall
  bowl (hemisphere, radius 0.08 m, two lattice layers) on its floor and a liquid block of U{2,000..20,000}
  particles above it, the worlds' particle counts uneven.
- ``pour_worlds(worlds, seed)``: batched pour worlds, each a 0.32 m box tank with a particle-wall bowl that sways
  along x (amplitude 0.05 m, period 1 s, phase U(0, 2 pi)), a spout at z = 0.25 m (half extents 0.01 m: a 2 x 2
  layer, 2 m/s down, a layer every 5 ms) and a pool of U{1,000..4,000} spares: the liquid grows from nothing.
Particle radius 0.005 m (h = 0.02 m, cubic kernel), rest density 1000 kg/m^3, XSPH 0.01, density and divergence
tolerance 0.1 %, resting-jitter rule off; dt = 1 ms is the caller's.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .build import add_box_tank, add_fluid_block, add_spares, add_wall_particles, bowl_points
from .config import DFSPHConfig
from .model import Model, ModelBuilder

RADIUS = 0.005
DT = 1.0e-3


def bench_config(**overrides) -> DFSPHConfig:
    """DFSPHConfig of the benchmark scenes: r = 5 mm, h = 2 cm, 0.1 % tolerances, resting-jitter rule off."""
    cfg = DFSPHConfig(**{"particle_radius": RADIUS, "rest_jitter_speed": None, "xsph_c": 0.01,
                         "viscosity": "xsph", **overrides})
    cfg.validate()
    return cfg


@dataclass
class Scene:
    model: Model
    config: DFSPHConfig
    inner_lo: list = field(default_factory=list)
    inner_hi: list = field(default_factory=list)
    fluid_counts: list = field(default_factory=list)
    emitters: list = field(default_factory=list)  # add_emitter arguments (world, region, half, rate, velocity)
    sway: dict | None = None                      # moving particle-wall bowls: wall ids, world phase, amplitude, period

    def make_solver(self, **kwargs):
        from .solver import SolverDFSPH

        s = SolverDFSPH(self.model, self.config, **kwargs)
        for args in self.emitters:
            s.add_emitter(*args)
        return s

    def drive(self, state, t: float) -> None:
        """Set the swaying bowls' wall-particle velocities for the step starting at time ``t``: v_x = A w cos(w t +
        phase_w), the particles then advance with it inside the step (particle_walls.py:131-135)."""
        if self.sway is None:
            return
        import torch

        sw = self.sway
        om = 2.0 * np.pi / sw["period"]
        vx = sw["amplitude"] * om * torch.cos(om * t + sw["phase"])
        state.particle_qd[sw["ids"], 0] = vx.to(state.particle_qd.dtype)


def dam_break(n_side: int = 20, box: float = 0.4, seed: int = 0, cfg: DFSPHConfig | None = None,
              device: str = "cpu", jitter: float = 1.0e-4, counts: tuple[int, int, int] | None = None) -> Scene:
    cfg = cfg or bench_config()
    rng = np.random.default_rng(seed)
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    b.begin_world()
    lo, hi = (0.0, 0.0, 0.0), (box, box, box)
    add_box_tank(b, lo, hi, wall=cfg.h)
    c = counts or (n_side, n_side, n_side)
    add_fluid_block(b, cfg, lo, c, jitter=jitter, rng=rng)
    b.end_world()
    return Scene(b.finalize(device), cfg, [lo], [hi], [int(np.prod(c))])


def water_column(nx: int = 6, ny: int = 6, nz: int = 20, *, worlds: int = 1, fills: list[int] | None = None,
                 cfg: DFSPHConfig | None = None, headroom: int = 8, spares: int = 0, device: str = "cpu") -> Scene:
    """A resting nx x ny x nz column in a box tank (walls h thick) per world, plus ``spares`` inactive spares per
    world for emitters (testing/sph.py:49-73)."""
    cfg = cfg or bench_config()
    fills = fills or [nz] * worlds
    s = cfg.spacing
    lo, hi = (0.0, 0.0, 0.0), (nx * s, ny * s, (max(fills) + headroom) * s)
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    for f in fills:
        b.begin_world()
        add_box_tank(b, lo, hi, wall=cfg.h)
        add_fluid_block(b, cfg, lo, (nx, ny, f))
        if spares:
            add_spares(b, cfg, spares)
        b.end_world()
    return Scene(b.finalize(device), cfg, [lo] * len(fills), [hi] * len(fills), [nx * ny * f for f in fills])


def bowl_worlds(worlds: int = 512, seed: int = 2, n_min: int = 2000, n_max: int = 20000, side=(0.3, 0.6),
                bowl_radius: float = 0.08, cfg: DFSPHConfig | None = None, device: str = "cpu",
                subset=None) -> Scene:
    """``worlds`` box tanks of side U(side) with a two-layer particle-wall bowl each and a lattice block of
    U[n_min, n_max] liquid particles (full layers of the block's footprint, the top layer partial) above it, ids in
    the lattice order of build.py:26-31 (x slowest). ``subset``: build only these world indices, each exactly as in
    the full pack (the seeded draws of every world are consumed either way)."""
    from .build import add_fluid_points, lattice_points

    cfg = cfg or bench_config()
    rng = np.random.default_rng(seed)
    sp = cfg.spacing
    keep_w = None if subset is None else set(int(w) for w in subset)
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    los, his, counts = [], [], []
    for w in range(worlds):
        L = float(rng.uniform(*side))
        N = int(rng.integers(n_min, n_max + 1))
        jitter_seed = rng.integers(1 << 31)
        if keep_w is not None and w not in keep_w:
            continue
        nxy = max(1, int(np.floor(0.9 * L / sp)))
        nz = int(np.ceil(N / (nxy * nxy)))
        rim = 2.0 * sp + bowl_radius
        z0 = rim + 2.0 * sp
        height = z0 + nz * sp + 0.1
        lo, hi = (0.0, 0.0, 0.0), (L, L, height)
        b.begin_world()
        add_box_tank(b, lo, hi, wall=cfg.h)
        add_wall_particles(b, cfg, bowl_points((0.5 * L, 0.5 * L, rim), bowl_radius, sp, layers=2))
        off = 0.5 * (L - nxy * sp)
        pts = lattice_points((off, off, z0), (nxy, nxy, nz), sp)
        pts = pts[np.lexsort((pts[:, 1], pts[:, 0], pts[:, 2]))][:N]  # the lowest N lattice sites, layer by layer
        pts = pts[np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0]))]       # back to x-slowest lattice order
        pts = pts + np.random.default_rng(jitter_seed).uniform(-1e-4, 1e-4, pts.shape)
        add_fluid_points(b, cfg, pts)
        b.end_world()
        los.append(lo)
        his.append(hi)
        counts.append(N)
    return Scene(b.finalize(device), cfg, los, his, counts)


def bowl_pool(worlds: int = 1, bowl_radius: float = 0.08, side: float = 0.3, cfg: DFSPHConfig | None = None,
              device: str = "cpu", seed: int = 0) -> Scene:
    """Liquid filling a particle-wall bowl (hemisphere of ``bowl_radius``, two lattice layers) in a box tank: every
    liquid particle near the shell reads the particle wall's colour field. Bowl radius grows 5 mm per world."""
    from .build import add_fluid_points, lattice_points

    cfg = cfg or bench_config()
    sp = cfg.spacing
    rng = np.random.default_rng(seed)
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    los, his, counts = [], [], []
    for w in range(worlds):
        R = bowl_radius + 0.005 * w
        rim = 2.0 * sp + R
        c = np.array([0.5 * side, 0.5 * side, rim])
        lo, hi = (0.0, 0.0, 0.0), (side, side, side)
        b.begin_world()
        add_box_tank(b, lo, hi, wall=cfg.h)
        add_wall_particles(b, cfg, bowl_points(tuple(c), R, sp, layers=2))
        m = int(np.ceil(R / sp)) + 1
        pts = lattice_points(tuple(c - m * sp), (2 * m, 2 * m, m), sp)
        keep = (np.linalg.norm(pts - c, axis=1) < R - 0.5 * sp) & (pts[:, 2] < c[2])
        pts = pts[keep] + rng.uniform(-1e-4, 1e-4, (int(keep.sum()), 3))
        add_fluid_points(b, cfg, pts)
        b.end_world()
        los.append(lo)
        his.append(hi)
        counts.append(len(pts))
    return Scene(b.finalize(device), cfg, los, his, counts)


def pour_worlds(worlds: int = 2048, seed: int = 4, cap=(1000, 4000), side: float = 0.32, bowl_radius: float = 0.08,
                spout_z: float = 0.25, spout_half: float = 0.01, spout_speed: float = 2.0, amplitude: float = 0.05,
                period: float = 1.0, cfg: DFSPHConfig | None = None, device: str = "cpu", subset=None) -> Scene:
    """Batched pour worlds: a box tank of side ``side``, a two-layer particle-wall bowl (radius ``bowl_radius``) on
    its floor swaying along x with ``amplitude`` and ``period`` at a phase U(0, 2 pi) per world, a box emitter at
    (side/2, side/2, spout_z) with half extents (spout_half, spout_half, 0.4 spacing) pouring down at ``spout_speed``
    at rate = layer area x speed (one layer per spacing / speed seconds), and U{cap} spares per world parked below
    the floor. Escape box: the tank plus 0.1 m. ``subset``: build only these worlds, each as in the full pack."""
    cfg = cfg or bench_config(world_bounds=((-0.1, -0.1, -0.1), (side + 0.1, side + 0.1, side + 0.1)))
    rng = np.random.default_rng(seed)
    sp = cfg.spacing
    keep_w = None if subset is None else set(int(w) for w in subset)
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    los, his, caps, phases, emit, wall_ids = [], [], [], [], [], []
    rim = 2.0 * sp + bowl_radius
    lo, hi = (0.0, 0.0, 0.0), (side, side, side)
    shell = bowl_points((0.5 * side, 0.5 * side, rim), bowl_radius, sp, layers=2)
    n_layer = max(int(np.floor(2.0 * spout_half / sp + 1.0e-6)), 1) ** 2
    rate = n_layer * sp * sp * spout_speed
    for w in range(worlds):
        n_cap = int(rng.integers(cap[0], cap[1] + 1))
        ph = float(rng.uniform(0.0, 2.0 * np.pi))
        if keep_w is not None and w not in keep_w:
            continue
        wl = len(los)
        b.begin_world()
        add_box_tank(b, lo, hi, wall=cfg.h)
        x0 = np.array([amplitude * np.sin(ph), 0.0, 0.0])  # the bowl's offset at t = 0
        r = add_wall_particles(b, cfg, shell + x0)
        wall_ids.append(np.arange(r.start, r.stop))
        add_spares(b, cfg, n_cap, park=(0.5 * side, 0.5 * side, -0.05))
        b.end_world()
        emit.append((wl, (0.5 * side, 0.5 * side, spout_z, 0.0, 0.0, 0.0, 1.0), (spout_half, spout_half, 0.4 * sp),
                     rate, (0.0, 0.0, -spout_speed)))
        los.append(lo)
        his.append(hi)
        caps.append(n_cap)
        phases.append(ph)
    model = b.finalize(device)
    import torch

    ids = torch.as_tensor(np.concatenate(wall_ids), dtype=torch.int64, device=device)
    ph_t = torch.as_tensor(np.repeat(phases, [len(x) for x in wall_ids]), dtype=torch.float64, device=device)
    sway = {"ids": ids, "phase": ph_t, "amplitude": amplitude, "period": period}
    return Scene(model, cfg, los, his, caps, emitters=emit, sway=sway)


__all__ = ["pour_worlds", "bowl_pool", "DT", "RADIUS", "Scene", "bench_config", "bowl_worlds", "dam_break", "water_column", "add_fluid_block"]
