"""Host-side (float64 numpy) smoothing-kernel constants of the DFSPH liquid.

These are the construction-time numbers the device kernels are parameterised with: the lattice kernel sum that
fixes the particle mass, the rest-lattice DFSPH diagonal, the fitted wall offset/amplitude and the neighbour count
of an interior lattice site. Reference: alloy/solvers/sph/kernels.py:155-282 (host mirrors) and :41-138 (the device
formulas these mirror). The device formulas themselves live in alloy_amd/csrc/dfsph_device.h.

Cubic spline with support h (W = 0 for r >= h): W = 8/(pi h^3) f(q), f = 6q^3 - 6q^2 + 1 for q <= 1/2 and
2(1-q)^3 for q <= 1. Wendland C2: 21/(2 pi h^3) (1-q)^4 (1+4q).

The flat-wall boundary volume Psi(d) is the kernel mass inside the solid half-space at signed distance d:
Psi(d) = 1/2 + sign(z) G0(|z|), z = clamp(-d/h, -1, 1), G0(y) = int_0^y p(t) dt with p the projected kernel
(kernels.py:80-112). The G0 polynomials below are integrals of the projected kernels and are checked against
quadrature in tests/test_kernels.py.
"""

from __future__ import annotations

import math

import numpy as np

KERNEL_IDS = {"cubic": 0, "wendland_c2": 1}


def kernel_w_np(r: np.ndarray, h: float, kind: str = "cubic") -> np.ndarray:
    """W(r) [1/m^3] in float64 (kernels.py:155-170)."""
    q = np.asarray(r, dtype=np.float64) / h
    w = np.zeros_like(q)
    norm3 = h**3
    if kind == "wendland_c2":
        inside = q < 1.0
        t = 1.0 - q[inside]
        w[inside] = 21.0 / (2.0 * math.pi * norm3) * t**4 * (1.0 + 4.0 * q[inside])
        return w
    c = 8.0 / (math.pi * norm3)
    near = q <= 0.5
    far = (q > 0.5) & (q < 1.0)
    w[near] = c * (6.0 * q[near] ** 3 - 6.0 * q[near] ** 2 + 1.0)
    w[far] = c * 2.0 * (1.0 - q[far]) ** 3
    return w


def kernel_dwdr_np(r: np.ndarray, h: float, kind: str = "cubic") -> np.ndarray:
    """dW/dr [1/m^4] in float64, the derivative of ``kernel_w_np``'s polynomials: cubic 8/(pi h^4) (18 q^2 - 12 q) for
    q <= 1/2 and -6 (1 - q)^2 8/(pi h^4) for q < 1; Wendland C2 -20 q (1 - q)^3 21/(2 pi h^4); 0 from q = 1."""
    q = np.asarray(r, dtype=np.float64) / h
    out = np.zeros_like(q)
    if kind == "wendland_c2":
        m = q < 1.0
        out[m] = 21.0 / (2.0 * math.pi * h**4) * (-20.0 * q[m] * (1.0 - q[m]) ** 3)
        return out
    c = 8.0 / (math.pi * h**4)
    near, far = q <= 0.5, (q > 0.5) & (q < 1.0)
    out[near] = c * (18.0 * q[near] ** 2 - 12.0 * q[near])
    out[far] = c * (-6.0 * (1.0 - q[far]) ** 2)
    return out


def _lattice_offsets(spacing: float, h: float) -> np.ndarray:
    """Offsets of a cubic lattice (spacing) covering the support ball, origin included, shape (M, 3)."""
    m = int(math.ceil(h / spacing)) + 1
    ax = np.arange(-m, m + 1, dtype=np.float64) * spacing
    return np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)


def lattice_sum(spacing: float, h: float, kind: str = "cubic") -> float:
    """S = sum of W over an infinite cubic lattice, self included [1/m^3] (kernels.py:173-179); m = rho0 / S."""
    g = _lattice_offsets(spacing, h)
    return float(kernel_w_np(np.linalg.norm(g, axis=1), h, kind).sum())


# G0(y) = int_0^y p on [0, top] as ascending polynomial coefficients (kernels.py:93-101, 182-186)
_G0_PIECES = {
    "cubic": ((0.5, (0.0, 1.4, 0.0, -8.0 / 3.0, 0.0, 4.8, -3.2)),
              (1.0, (-1.0 / 30.0, 1.6, 0.0, -16.0 / 3.0, 8.0, -4.8, 16.0 / 15.0))),
    "wendland_c2": ((1.0, (0.0, 1.5, 0.0, -3.5, 0.0, 10.5, -14.0, 7.5, -1.5)),),
}


def _g0_np(y: float, kind: str) -> float:
    for top, coeffs in _G0_PIECES[kind]:
        if y <= top:
            return float(np.polynomial.polynomial.polyval(y, coeffs))
    return 0.5


def halfspace_volume_np(d: float, h: float, kind: str = "cubic") -> float:
    """Psi(d) of a flat wall (H = 0) in float64 (kernels.py:196-201)."""
    if d >= h:
        return 0.0
    z = min(max(-d / h, -1.0), 1.0)
    return 0.5 + math.copysign(_g0_np(abs(z), kind), z)


def _half_lattice(spacing: float, h: float, kind: str) -> tuple[float, float]:
    """(1 - F1, |dF1/dz|) for a particle on the first lattice layer beside a flat wall (kernels.py:204-218): F1 is
    the kernel mass of the layers z >= 0 normalised by the full lattice sum."""
    g = _lattice_offsets(spacing, h)
    r = np.linalg.norm(g, axis=1)
    w = kernel_w_np(r, h, kind)
    total = float(w.sum())
    upper = (g[:, 2] >= 0.0) & (r > 0.0)
    e = 1.0e-7 * h
    dwdr = (kernel_w_np(r[upper] + e, h, kind) - kernel_w_np(r[upper] - e, h, kind)) / (2.0 * e)
    slope = float(np.sum(dwdr * (-g[upper, 2]) / r[upper])) / total
    f1 = float(w[g[:, 2] >= 0.0].sum()) / total
    return 1.0 - f1, abs(slope)


def boundary_fit(spacing: float, h: float, kind: str = "cubic", slope_factor: float = 2.0) -> tuple[float, float]:
    """(delta, A) of the fitted wall volume Psi_eff(d) = A Psi(d + delta) (kernels.py:221-249).

    At the rest layer d = r the fitted volume equals the half-lattice deficit (a lattice beside a flat wall sits at
    rho0) and its slope is ``slope_factor`` times the half-lattice gradient. Psi/|Psi'| falls monotonically with d,
    so the matching distance is found by bisection; delta is that distance minus r."""
    deficit, grad = _half_lattice(spacing, h, kind)
    target = deficit / (slope_factor * grad)

    def psi_over_slope(d: float) -> float:
        e = 1.0e-7 * h
        slope = (halfspace_volume_np(d + e, h, kind) - halfspace_volume_np(d - e, h, kind)) / (2.0 * e)
        return halfspace_volume_np(d, h, kind) / max(abs(slope), 1.0e-30)

    lo, hi = 0.0, 0.999 * h
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if psi_over_slope(mid) > target:
            lo = mid
        else:
            hi = mid
    d = 0.5 * (lo + hi)
    return d - 0.5 * spacing, deficit / halfspace_volume_np(d, h, kind)


def _lattice_dwdr(spacing: float, h: float, kind: str) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    g = _lattice_offsets(spacing, h)
    r = np.linalg.norm(g, axis=1)
    sel = (r > 0.0) & (r < h)
    e = 1.0e-7 * h
    dw = (kernel_w_np(r[sel] + e, h, kind) - kernel_w_np(r[sel] - e, h, kind)) / (2.0 * e)
    return g[sel], r[sel], dw


LAPLACE_D = 10.0  # 2 (dim + 2), the Weiler/Monaghan viscous Laplacian's coefficient (viscosity.py:34)


def laplacian_lattice_factor(spacing: float, h: float, kind: str = "cubic") -> float:
    """Discrete / exact ratio of the viscous Laplacian 2 (d + 2) sum_j V (v_ij . x_ij) / (|x_ij|^2 + 0.01 h^2) grad W_ij
    on the rest lattice (V = 1 / S) for the shear v = (z^2 / 2, 0, 0), whose exact Laplacian is 1 (kernels.py:252-262):
    the implicit viscosity divides its coefficient by it, so the configured viscosity is the effective one."""
    g, r, dw = _lattice_dwdr(spacing, h, kind)
    grad_x = dw * g[:, 0] / r
    v_ij_x = -0.5 * g[:, 2] ** 2
    vol = 1.0 / lattice_sum(spacing, h, kind)
    return float(LAPLACE_D * vol * np.sum(v_ij_x * g[:, 0] / (r**2 + 0.01 * h * h) * grad_x))


def lattice_diagonal(spacing: float, h: float, kind: str = "cubic") -> float:
    """Interior rest-lattice DFSPH diagonal sum_j (V |dW_ij|)^2 with V = 1/S [1/m^2] (kernels.py:265-270)."""
    _g, _r, dw = _lattice_dwdr(spacing, h, kind)
    v = 1.0 / lattice_sum(spacing, h, kind)
    return float(np.sum((v * dw) ** 2))


def full_neighbor_count(spacing: float, h: float, kind: str = "cubic") -> float:
    """Neighbours within h of an interior lattice site, self excluded (config.py:230-236)."""
    g = _lattice_offsets(spacing, h)
    return float(np.count_nonzero(kernel_w_np(np.linalg.norm(g, axis=1), h, kind) > 0.0) - 1)
