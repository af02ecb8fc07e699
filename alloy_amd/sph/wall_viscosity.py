"""Wall (boundary) viscosity of the implicit viscosity solve: constants and the coefficient calibration
(wall_viscosity.py:1-93 of the reference; the per-slot terms run in the HIP library, ``k_visc_wall_terms``).

For each boundary slot of a liquid particle the term couples the particle to four samples around the slot's
representative boundary point x_k = x_i - d n (d = max(phi + r/2, 2r)), offset by +-TANGENTIAL_FACTOR h along two
tangents; the normal block comes from the samples, the tangential block is the no-slip stress -mu_b / (max(phi, r)
2r), faded out from TANGENT_FULL r to TANGENT_ZERO r. ``wall_coefficient_scale`` is the per-unit-mu_b coefficient
that makes a first-layer particle of the rest lattice feel exactly the no-slip stress mu_b u / r on its (2r)^2 face.
"""

from __future__ import annotations

import math

import numpy as np

from .kernels import _half_lattice, kernel_dwdr_np, kernel_w_np

TANGENTIAL_FACTOR = 0.5  # sample offset along the wall, in h (wall_viscosity.py:51)
TANGENT_FULL, TANGENT_ZERO = 2.0, 3.0  # in r: full tangential block up to 2r, none from 3r (wall_viscosity.py:52)


def wall_coefficient_scale(spacing: float, h: float, kind: str = "cubic") -> float:
    """Coefficient per unit mu_b (coeff = mu_b * scale) giving a first-layer particle of the rest lattice (phi = r,
    boundary volume = the lattice deficit) the tangential wall block -mu_b / (2 r^2) (wall_viscosity.py:75-93).

    The slot's boundary point lies max(phi + r/2, 2r) below the particle; the lattice deficit Psi it stands for is
    shared by its four samples, each of volume Psi / (4 W(depth)). The two samples offset by TANGENTIAL_FACTOR h
    along the shear give the tangential block per unit coefficient, 2 V W'(s) / s * o^2 / (s^2 + 0.01 h^2) at the
    sample distance s = |(depth, o)|, with W' the kernel's analytic slope."""
    r = 0.5 * spacing
    depth = max(r + 0.5 * r, 2.0 * r)
    deficit = _half_lattice(spacing, h, kind)[0]
    volume = deficit / (4.0 * float(kernel_w_np(np.array([depth]), h, kind)[0]))
    o = TANGENTIAL_FACTOR * h
    s = math.hypot(depth, o)
    slope = float(kernel_dwdr_np(np.array([s]), h, kind)[0])
    block = 2.0 * volume * (slope / s) * (o * o) / (s * s + 0.01 * h * h)
    if not block < 0.0:
        raise ValueError(f"wall_coefficient_scale: no wall term reaches the first layer (h={h}, spacing={spacing})")
    return 1.0 / (2.0 * r * r * -block)
