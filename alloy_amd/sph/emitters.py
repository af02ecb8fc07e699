"""Box emitters, sinks and the escape box of SolverDFSPH: the host bookkeeping.

Follows alloy/solvers/sph/emitters.py:32-64 and 175-255 (EmitterTable, layer_offsets, split_pools, rewind) and
runtime.py:197-304 (EmitterManager.add_emitter / add_sink / _split_pools / _upload). The device work (the clock, the
layer emission, sinks, the escape box and the int64 ledgers) is the emitter pass of libdfsph_hip.so, which
``dfsph_step`` runs after advection when any emitter, sink or escape box exists.

An emitter owns a pool of its world's spare particles (sph:phase 0, ACTIVE cleared): the ids it was given, else an
even share of the spares no emitter owns. Every ``interval = n_layer V / rate`` seconds it activates the next layer
of its pool on a 2r lattice over the region's cross-section perpendicular to the velocity. Pools never refill.
"""

from __future__ import annotations

import ctypes
import logging
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from .config import PHASE_HELD, PHASE_SPARE
from .model import ParticleFlags

log = logging.getLogger("alloy.sph")


@dataclass
class EmitterTable:
    """Host-side emitter and sink definitions (emitters.py:32-44); uploaded to the device on every change."""

    world: list[int] = field(default_factory=list)
    xform: list[tuple] = field(default_factory=list)  # p(3), q(4, xyzw)
    velocity: list[tuple[float, float, float]] = field(default_factory=list)
    interval: list[float] = field(default_factory=list)
    offsets: list[np.ndarray] = field(default_factory=list)  # layer lattice per emitter, region frame
    pool: list[np.ndarray] = field(default_factory=list)  # spare particle ids per emitter, emission order
    own: list[bool] = field(default_factory=list)  # True: the pool was given at add_emitter
    sink_world: list[int] = field(default_factory=list)
    sink_xform: list[tuple] = field(default_factory=list)
    sink_half: list[tuple[float, float, float]] = field(default_factory=list)


def as_transform(region) -> tuple:
    """A transform as 7 floats p(3), q(4, xyzw): a 7-sequence, or a (position, quaternion) pair."""
    if len(region) == 2:
        p, q = region
        vals = [*p, *q]
    else:
        vals = list(region)
    if len(vals) != 7:
        raise ValueError(f"a transform is p(3) + q(4, xyzw), got {len(vals)} values")
    return tuple(float(v) for v in vals)


def quat_rotate_inv(q, v) -> np.ndarray:
    """v rotated by the inverse of the unit quaternion q (x, y, z, w), float64."""
    u = -np.asarray(q[:3], np.float64)
    w = float(q[3])
    v = np.asarray(v, np.float64)
    return v * (2.0 * w * w - 1.0) + np.cross(u, v) * w * 2.0 + u * np.dot(u, v) * 2.0


def layer_offsets(half_extents, velocity_local, spacing: float) -> np.ndarray:
    """Lattice points (region frame) on the region's mid cross-section perpendicular to the velocity's dominant
    axis: floor(2 half / spacing + 1e-6) columns (at least 1) per side, centred (emitters.py:46-64)."""
    v = np.asarray(velocity_local, dtype=np.float64)
    half = np.asarray(half_extents, dtype=np.float64)
    axis = int(np.argmax(np.abs(v))) if float(np.linalg.norm(v)) > 0.0 else 2
    others = [a for a in range(3) if a != axis]
    cols = []
    for a in others:
        k = max(int(np.floor(2.0 * half[a] / spacing + 1.0e-6)), 1)
        cols.append((np.arange(k) - 0.5 * (k - 1)) * spacing)
    ga, gb = np.meshgrid(cols[0], cols[1], indexing="ij")
    pts = np.zeros((ga.size, 3))
    pts[:, others[0]] = ga.ravel()
    pts[:, others[1]] = gb.ravel()
    return pts.astype(np.float32)


def spare_ids(phase: np.ndarray, world: np.ndarray, flags: np.ndarray, w: int) -> np.ndarray:
    """Inactive SPH spares of world w, ascending ids (layout.py:279-282)."""
    mask = (phase == PHASE_SPARE) & (world == w) & ((flags & int(ParticleFlags.ACTIVE)) == 0)
    return np.nonzero(mask)[0].astype(np.int32)


def split_pools(t: EmitterTable, world: int, spares: np.ndarray, cursors: np.ndarray) -> None:
    """Share ``world``'s spares evenly among its emitters that were given no pool; an emitter given its pool keeps it.
    Once one of them has emitted, only the last one's share is recomputed (emitters.py:215-231)."""
    owned = [t.pool[e] for e, w in enumerate(t.world) if w == world and t.own[e]]
    free = np.setdiff1d(spares, np.concatenate(owned)) if owned else spares
    mine = [e for e, w in enumerate(t.world) if w == world and not t.own[e]]
    if any(e < len(cursors) and cursors[e] > 0 for e in mine[:-1]):
        used = np.concatenate([t.pool[e] for e in mine[:-1]]) if len(mine) > 1 else np.zeros(0, np.int32)
        t.pool[mine[-1]] = np.setdiff1d(free, used).astype(np.int32)
        return
    for k, part in enumerate(np.array_split(free, len(mine)) if mine else ()):
        t.pool[mine[k]] = part.astype(np.int32)


class EmitterStruct:
    """The DfsphEmitters struct of a table (include/dfsph_abi.h), with the host arrays it points to kept alive."""

    def __init__(self, t: EmitterTable, bounds=None):
        E, K = len(t.world), len(t.sink_world)
        i32 = lambda a: np.ascontiguousarray(np.asarray(a, np.int32).reshape(-1))  # noqa: E731
        f32 = lambda a: np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1))  # noqa: E731
        counts = [len(o) for o in t.offsets]
        sizes = [len(p) for p in t.pool]
        self.keep = dict(
            world=i32(t.world), xform=f32(t.xform), velocity=f32(t.velocity), interval=f32(t.interval),
            offset_start=i32(np.concatenate([[0], np.cumsum(counts)])) if E else i32([0]),
            offsets=f32(np.concatenate(t.offsets)) if E and sum(counts) else f32([0.0]),
            pool_start=i32(np.concatenate([[0], np.cumsum(sizes)])) if E else i32([0]),
            pool_ids=i32(np.concatenate(t.pool)) if E and sum(sizes) else i32([0]),
            sink_world=i32(t.sink_world), sink_xform=f32(t.sink_xform), sink_half=f32(t.sink_half))
        s = _abi.DfsphEmitters()
        s.emitter_count, s.sink_count = E, K
        for name in ("world", "offset_start", "pool_start", "pool_ids", "sink_world"):
            setattr(s, name, self.keep[name].ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        for name in ("xform", "velocity", "interval", "offsets", "sink_xform", "sink_half"):
            setattr(s, name, self.keep[name].ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        s.has_bounds = int(bounds is not None)
        if bounds is not None:
            for k in range(3):
                s.bounds_lo[k] = float(bounds[0][k])
                s.bounds_hi[k] = float(bounds[1][k])
        self.struct = s


class EmitterManager:
    """Host bookkeeping of one solver's emitters and sinks (runtime.py:197-304); ``upload(struct)`` hands the table
    to the device (the library, or the oracle in tests). With ``eager`` False a change marks the table dirty and
    ``flush`` uploads it (the solver flushes before a step or a read of emitter state), so adding thousands of
    emitters uploads once."""

    def __init__(self, model, config, upload, cursors, eager: bool = True):
        self.model, self.config = model, config
        self._upload_fn = upload  # callable(EmitterStruct)
        self._cursors = cursors   # callable() -> np.ndarray of the device cursors (or zeros)
        self.eager = eager
        self.dirty = False
        self.table = EmitterTable()
        b = config.world_bounds
        self.bounds = None if b is None else (tuple(float(v) for v in b[0]), tuple(float(v) for v in b[1]))
        self.struct: EmitterStruct | None = None

    @property
    def active(self) -> bool:
        return bool(self.table.world) or bool(self.table.sink_world) or self.bounds is not None

    def _spares(self, w: int) -> np.ndarray:
        """Inactive spares of world w, ascending ids: only the world's own slice crosses to the host."""
        m = self.model
        a, b = int(m.world_start_np[w]), int(m.world_start_np[w + 1])
        ph = m.sph.phase[a:b].cpu().numpy()
        fl = m.particle_flags[a:b].cpu().numpy()
        return (a + spare_ids(ph, np.full(b - a, w), fl, w)).astype(np.int32)

    def _first_sph(self, world: int) -> int:
        ws = self.model.world_start_np
        ph = self.model.sph.phase[ws[world]:ws[world + 1]].cpu().numpy()
        idx = np.nonzero((ph >= PHASE_SPARE) | (ph == PHASE_HELD))[0]
        if len(idx) == 0:
            raise ValueError(f"world {world} has no SPH particles or spares")
        return int(ws[world] + idx[0])

    def add_emitter(self, world: int, region, half_extents, rate: float, velocity, pool=None) -> int:
        """runtime.py:243-270"""
        if not 0 <= world < self.model.world_count:
            raise ValueError(f"add_emitter: world {world} out of range")
        if not rate > 0.0:
            raise ValueError("add_emitter: rate must be > 0 [m^3/s]")
        X = as_transform(region)
        vel = tuple(float(c) for c in velocity)
        v_local = quat_rotate_inv(X[3:], vel)
        offs = layer_offsets(np.asarray(half_extents, dtype=np.float64), v_local, self.config.spacing)
        first = self._first_sph(world)
        rho0 = float(self.model.sph.rest_density[first].item())
        mass = float(self.model.particle_mass[first].item())
        interval = len(offs) * (mass / rho0) / rate
        self._check_rate(offs, v_local, rate, self.config.spacing)
        self._check_inside_bounds(np.asarray(X[:3], dtype=np.float64))
        t = self.table
        t.world.append(int(world))
        t.xform.append(X)
        t.velocity.append(vel)
        t.interval.append(interval)
        t.offsets.append(offs)
        t.own.append(pool is not None)
        t.pool.append(np.zeros(0, np.int32) if pool is None
                      else np.intersect1d(np.asarray(pool), self._spares(world)).astype(np.int32))
        split_pools(t, world, self._spares(world), self._cursors())
        self.upload()
        log.info("SPH emitter %d: world %d, %d particles per layer every %.4g s", len(t.world) - 1, world,
                 len(offs), interval)
        return len(t.world) - 1

    @staticmethod
    def _check_rate(offs: np.ndarray, v_local: np.ndarray, rate: float, spacing: float) -> None:
        area = len(offs) * spacing * spacing
        speed = float(np.linalg.norm(v_local))
        if speed > 0.0 and not 0.8 <= rate / (area * speed) <= 1.25:
            log.warning("SPH emitter rate %.3g m^3/s vs area*|v| %.3g m^3/s: layers will over/under-pack", rate,
                        area * speed)

    def _check_inside_bounds(self, centre: np.ndarray) -> None:
        if self.bounds is None:
            return
        lo, hi = np.array(self.bounds[0]), np.array(self.bounds[1])
        if np.any(centre < lo) or np.any(centre > hi):
            log.warning("SPH add_emitter: region centre %s is outside world_bounds %s; its liquid retires as it is "
                        "emitted (report().escaped)", centre.tolist(), (lo.tolist(), hi.tolist()))

    def add_sink(self, world: int, region, half_extents) -> int:
        """runtime.py:286-294"""
        if not 0 <= world < self.model.world_count:
            raise ValueError(f"add_sink: world {world} out of range")
        self.table.sink_world.append(int(world))
        self.table.sink_xform.append(as_transform(region))
        self.table.sink_half.append(tuple(float(c) for c in half_extents))
        self.upload()
        return len(self.table.sink_world) - 1

    def upload(self) -> None:
        if not self.eager:
            self.dirty = True
            return
        self.flush(force=True)

    def flush(self, force: bool = False) -> None:
        if not (force or self.dirty):
            return
        self.struct = EmitterStruct(self.table, self.bounds)
        self._upload_fn(self.struct)
        self.dirty = False

    def rewound_emitters(self, world_mask: np.ndarray | None) -> np.ndarray:
        """Indices of the emitters in the masked worlds (all when None): ``emitters.rewind`` sets their clock and
        cursor back to 0 (emitters.py:234-255)."""
        worlds = np.asarray(self.table.world, np.int64)
        if world_mask is None:
            return np.arange(len(worlds))
        mask = np.asarray(world_mask, bool)
        return np.nonzero(mask[worlds])[0] if len(worlds) else np.zeros(0, np.int64)
