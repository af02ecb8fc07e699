"""The slice of Newton's Model/State/ModelBuilder the DFSPH plug-in reads, on torch tensors.

SolverDFSPH reads a Newton ``Model`` through these field names (alloy/solvers/sph/solver.py:188-196,
layout.py:47-72, boundary.py:79-95): ``particle_q/qd/mass/flags/world/radius``, ``particle_world_start``,
``gravity[W]``, ``sph.rest_density/phase`` and the particle-colliding shapes ``shape_type/scale/transform/body/
world/flags/material_mu/collision_radius``. This module provides exactly those fields, built on the host and
finalised to torch tensors on one device, so scenes and the solver need neither Warp nor Newton.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from enum import IntEnum

import numpy as np
import torch


class ParticleFlags(IntEnum):
    """Newton's particle flags (alloy/_vendor/newton/_src/geometry/flags.py:9-24)."""

    ACTIVE = 1 << 0
    PROXY = 1 << 1


class ModelFlags(IntEnum):
    """Newton's notify_model_changed flags (alloy/_vendor/newton/_src/sim/enums.py:8-38), the ones this path reads."""

    BODY_PROPERTIES = 1 << 2
    BODY_INERTIAL_PROPERTIES = 1 << 3
    SHAPE_PROPERTIES = 1 << 4
    MODEL_PROPERTIES = 1 << 5


class ShapeFlags(IntEnum):
    VISIBLE = 1 << 0
    COLLIDE_SHAPES = 1 << 1
    COLLIDE_PARTICLES = 1 << 2


class GeoType(IntEnum):
    """Newton's GeoType values (alloy/_vendor/newton/_src/geometry/types.py:73-100) for the shapes handled here."""

    PLANE = 1
    SPHERE = 3
    BOX = 7


@dataclass
class SPHAttributes:
    """model.sph: per-particle rest density [kg/m^3], viscosity [Pa s], phase, surface tension sigma [N/m] and
    Akinci adhesion, and per shape the contact angle [rad] (-1: a free contact line) (config.py:244-261)."""

    rest_density: torch.Tensor
    viscosity: torch.Tensor
    phase: torch.Tensor
    surface_tension: torch.Tensor | None = None  # [n] float32, 0: no tension
    adhesion: torch.Tensor | None = None         # [n] float32 (read by akinci2013 only)
    shape_contact_angle: np.ndarray | None = None  # [S] float32 host array, -1: free

    def __post_init__(self):
        n = self.rest_density.shape[0]
        dev = self.rest_density.device
        if self.surface_tension is None:
            self.surface_tension = torch.zeros(n, dtype=torch.float32, device=dev)
        if self.adhesion is None:
            self.adhesion = torch.zeros(n, dtype=torch.float32, device=dev)
        if self.shape_contact_angle is None:
            self.shape_contact_angle = np.zeros(0, np.float32)


@dataclass
class State:
    particle_q: torch.Tensor   # [n, 3] float32
    particle_qd: torch.Tensor  # [n, 3] float32
    particle_f: torch.Tensor | None = None


@dataclass
class Model:
    device: torch.device
    particle_count: int
    world_count: int
    particle_q: torch.Tensor
    particle_qd: torch.Tensor
    particle_mass: torch.Tensor
    particle_radius: torch.Tensor
    particle_flags: torch.Tensor
    particle_world: torch.Tensor
    particle_world_start: torch.Tensor  # [W + 1] int32 (host mirror in world_start_np)
    gravity: torch.Tensor               # [W, 3]
    sph: SPHAttributes
    shape_count: int
    shape_type: np.ndarray
    shape_scale: np.ndarray
    shape_transform: np.ndarray         # [S, 7] p, q (xyzw)
    shape_body: np.ndarray
    shape_world: np.ndarray
    shape_flags: np.ndarray
    shape_material_mu: np.ndarray
    shape_collision_radius: np.ndarray
    world_start_np: np.ndarray = field(default_factory=lambda: np.zeros(1, np.int64))

    def state(self) -> State:
        """A new State holding copies of the model's initial particle positions and velocities."""
        return State(self.particle_q.clone(), self.particle_qd.clone(), None)


class ModelBuilder:
    """Accumulates worlds of particles and shapes on the host; ``finalize`` moves them to a device.

    Worlds are contiguous particle ranges (begin_world / end_world), as Newton's builder lays them out
    (layout.py:47-72 requires it)."""

    def __init__(self, gravity: tuple[float, float, float] = (0.0, 0.0, -9.81)):
        self.gravity = tuple(float(g) for g in gravity)
        self._q: list = []
        self._qd: list = []
        self._mass: list = []
        self._radius: list = []
        self._flags: list = []
        self._world: list = []
        self._rho0: list = []
        self._mu: list = []
        self._phase: list = []
        self._sigma: list = []
        self._adh: list = []
        self._shapes: list[dict] = []
        self._world_gravity: list = []
        self.current_world = -1
        self.world_count = 0
        self._sph_registered = False
        self._count = 0

    # ---------------------------------------------------------------------------------------- worlds
    def begin_world(self, gravity: tuple[float, float, float] | None = None) -> int:
        if self.current_world >= 0:
            raise ValueError("ModelBuilder: begin_world inside a world")
        self.current_world = self.world_count
        self.world_count += 1
        self._world_gravity.append(tuple(float(g) for g in (gravity or self.gravity)))
        return self.current_world

    def end_world(self) -> None:
        if self.current_world < 0:
            raise ValueError("ModelBuilder: end_world without begin_world")
        self.current_world = -1

    def register_sph_attributes(self) -> None:
        self._sph_registered = True

    @property
    def particle_count(self) -> int:
        return self._count

    # ---------------------------------------------------------------------------------------- content
    def add_particles(self, pos, vel, mass, radius, flags, rest_density=1000.0, viscosity=1.0e-3,
                      phase=-1, surface_tension=0.0, adhesion=0.0) -> range:
        if self.current_world < 0:
            raise ValueError("ModelBuilder: particles must be added inside a world")
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = len(pos)
        first = self.particle_count
        bc = lambda v, dt: np.broadcast_to(np.asarray(v, dtype=dt), (n,)).copy()  # noqa: E731
        self._q.append(pos)
        self._qd.append(np.broadcast_to(np.asarray(vel, dtype=np.float64), (n, 3)).copy())
        self._mass.append(bc(mass, np.float64))
        self._radius.append(bc(radius, np.float64))
        self._flags.append(bc(flags, np.int32))
        self._world.append(np.full(n, self.current_world, np.int32))
        self._rho0.append(bc(rest_density, np.float64))
        self._mu.append(bc(viscosity, np.float64))
        self._phase.append(bc(phase, np.int8))
        self._sigma.append(bc(surface_tension, np.float64))
        self._adh.append(bc(adhesion, np.float64))
        self._count += n
        return range(first, first + n)

    def add_shape(self, geo: int, center, half, quat=(0.0, 0.0, 0.0, 1.0), body: int = -1, mu: float = 0.5,
                  radius: float | None = None, contact_angle: float = -1.0) -> int:
        """A shape in the current world (or global with no world open). ``half``: box half extents, sphere
        radius in x, plane footprint (<= 0: infinite) in x/y; ``contact_angle`` [rad] for csf (-1: free)."""
        c = np.asarray(center, dtype=np.float64)
        hx = np.asarray(half, dtype=np.float64).reshape(3)
        if radius is None:
            radius = float(np.linalg.norm(hx)) if geo == GeoType.BOX else float(hx[0]) if geo == GeoType.SPHERE \
                else 1.0e30
        self._shapes.append(dict(geo=int(geo), center=c, half=hx, quat=np.asarray(quat, np.float64), body=int(body),
                                 world=self.current_world, mu=float(mu), radius=float(radius),
                                 contact_angle=float(contact_angle),
                                 flags=int(ShapeFlags.COLLIDE_PARTICLES | ShapeFlags.COLLIDE_SHAPES)))
        return len(self._shapes) - 1

    def add_shape_box(self, body: int, center, half, mu: float = 0.5) -> int:
        return self.add_shape(GeoType.BOX, center, half, body=body, mu=mu)

    # ---------------------------------------------------------------------------------------- finalize
    def finalize(self, device: str | torch.device = "cpu") -> Model:
        if self.current_world >= 0:
            raise ValueError("ModelBuilder: finalize inside a world")
        dev = torch.device(device)
        n = self.particle_count
        W = max(self.world_count, 1)
        cat = lambda parts, dt, shape=(0,): np.concatenate(parts).astype(dt) if parts else np.zeros(shape, dt)  # noqa: E731
        world = cat(self._world, np.int32)
        if n and np.any(np.diff(world) < 0):
            raise ValueError("particles must be contiguous per world")
        ws = np.zeros(W + 1, dtype=np.int64)
        ws[1:] = np.cumsum(np.bincount(world, minlength=W)[:W])
        f32 = lambda a: torch.tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)  # noqa: E731
        grav = np.asarray(self._world_gravity or [self.gravity], dtype=np.float32).reshape(-1, 3)
        S = len(self._shapes)
        sh = self._shapes
        xform = np.zeros((S, 7), np.float32)
        for s, d in enumerate(sh):
            xform[s, :3] = d["center"]
            xform[s, 3:] = d["quat"]
        model = Model(
            device=dev,
            particle_count=n,
            world_count=W,
            particle_q=f32(cat(self._q, np.float64, (0, 3)).reshape(-1, 3)),
            particle_qd=f32(cat(self._qd, np.float64, (0, 3)).reshape(-1, 3)),
            particle_mass=f32(cat(self._mass, np.float64)),
            particle_radius=f32(cat(self._radius, np.float64)),
            particle_flags=torch.tensor(cat(self._flags, np.int32), device=dev),
            particle_world=torch.tensor(world, device=dev),
            particle_world_start=torch.tensor(ws.astype(np.int32), device=dev),
            gravity=torch.tensor(grav, device=dev),
            sph=SPHAttributes(rest_density=f32(cat(self._rho0, np.float64)), viscosity=f32(cat(self._mu, np.float64)),
                              phase=torch.tensor(cat(self._phase, np.int8), device=dev),
                              surface_tension=f32(cat(self._sigma, np.float64)), adhesion=f32(cat(self._adh, np.float64)),
                              shape_contact_angle=np.asarray([d["contact_angle"] for d in sh], np.float32)),
            shape_count=S,
            shape_type=np.asarray([d["geo"] for d in sh], np.int32),
            shape_scale=np.asarray([d["half"] for d in sh], np.float32).reshape(S, 3),
            shape_transform=xform,
            shape_body=np.asarray([d["body"] for d in sh], np.int32),
            shape_world=np.asarray([d["world"] for d in sh], np.int32),
            shape_flags=np.asarray([d["flags"] for d in sh], np.int32),
            shape_material_mu=np.asarray([d["mu"] for d in sh], np.float32),
            shape_collision_radius=np.asarray([d["radius"] for d in sh], np.float32),
            world_start_np=ws,
        )
        return model
