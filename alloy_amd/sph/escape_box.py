"""The default escape box of a scene-built SolverDFSPH (alloy/solvers/sph/escape_box.py:1-94).

When the ``sph`` factory gets neither a ``config`` nor a ``world_bounds`` kwarg, ``world_bounds`` comes from the
scene's content: the box around the finite shapes (bounding spheres at their build pose; planes left out) and every
particle with a finite position (spares too), padded on every side by max(AUTO_PAD_MIN, largest extent). Liquid
that leaves it retires: its mass goes to the ledger's ``out`` and the count to ``SPHReport.escaped``. One box serves
every world. The drive-widening of the reference (``widen``, ``drive_reach``) applies to kinematic bodies, which
this solver does not own.
"""

from __future__ import annotations

from typing import Any

import numpy as np

AUTO_PAD_MIN = 0.25  # smallest padding [m] around the content box
_PLANE_RADIUS = 1.0e5  # planes report an unbounded collision radius

Bounds = tuple[tuple[float, float, float], tuple[float, float, float]]


def content_box(model: Any) -> tuple[np.ndarray, np.ndarray] | None:
    """Min and max corner of the finite static shapes and the particles over all worlds; None for an empty model."""
    pts = [np.zeros((0, 3))]
    lo_hi = []
    if int(getattr(model, "shape_count", 0) or 0):
        x = np.asarray(model.shape_transform, np.float64).reshape(-1, 7)
        r = np.asarray(model.shape_collision_radius, np.float64)
        keep = np.isfinite(r) & (r < _PLANE_RADIUS) & np.isfinite(x[:, :3]).all(axis=1)
        c, r = x[keep, :3], r[keep]
        if len(c):
            lo_hi.append((c - r[:, None], c + r[:, None]))
    if int(getattr(model, "particle_count", 0) or 0):
        q = model.particle_q.detach().cpu().numpy().astype(np.float64).reshape(-1, 3)
        pts.append(q[np.isfinite(q).all(axis=1)])
    p = np.vstack(pts)
    if not lo_hi and len(p) == 0:
        return None
    lo = np.vstack([a for a, _ in lo_hi] + [p]).min(axis=0)
    hi = np.vstack([b for _, b in lo_hi] + [p]).max(axis=0)
    return lo, hi


def auto_box(model: Any, pad_min: float = AUTO_PAD_MIN) -> tuple[Bounds, float] | None:
    """(The content box grown on every side by its largest extent, at least ``pad_min``; that growth), or None for
    an empty model."""
    box = content_box(model)
    if box is None:
        return None
    lo, hi = (np.asarray(c, np.float64) for c in box)
    grow = float(np.max(hi - lo))
    if grow < pad_min:
        grow = float(pad_min)
    corners = tuple(tuple(float(v) for v in c) for c in (lo - grow, hi + grow))
    return corners, grow


def auto_world_bounds(model: Any, pad_min: float = AUTO_PAD_MIN) -> Bounds | None:
    """The corners of ``auto_box``; None for an empty model."""
    box = auto_box(model, pad_min)
    return box[0] if box is not None else None
