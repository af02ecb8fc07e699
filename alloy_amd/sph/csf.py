"""Continuum surface force with a wall contact angle (``surface_tension="csf"``): the host side (csf.py,
csf_sparse.py and the calibration of tension.py in the reference). The per-step passes run in the HIP library
(``k_csf_*``), as one more stage of the non-pressure update before the implicit viscosity.

The method, per step (csf.py:1-36):
- colour gradient g_i = sum_j (m_j / rho0_i) grad W_ij + 1/2 sum_slots grad Psi (a wetted wall counts as liquid);
  |g| h > SURFACE_GRAD marks a surface particle, interior particles get no force;
- normal n = -<g> / |<g>| with <g> the gradient Shepard-smoothed NORMAL_PASSES + 1 times; at a surface particle whose
  nearest slot's shape has a contact angle theta >= 0, n turns to cos(theta) n_w + sin(theta) t (Breinlinger 2013);
- curvature kappa = the least-squares tangential divergence of n over the surface neighbours, Shepard-smoothed
  CURVATURE_PASSES times over the surface and moved from the particle centre to the interface;
- dv_i = -dt w_i sigma kappa |g| n / (rho_i I), I = ``flat_lattice_constants``' delta normalisation, w_i the CSF
  weight of the largest neighbour count around i; at a contact line the part pointing away from the wall goes to the
  wall's slot impulse;
- small pieces (no full neighbourhood nearby, CSF_BAND): Akinci's pair force (cohesion + curvature, without
  adhesion) with gamma from the Kirkwood-Buff calibration of the rest lattice, weighted by 1 - (w_i + w_j) / 2.
"""

from __future__ import annotations

import math

import numpy as np

from .kernels import kernel_w_np

SURFACE_GRAD = 0.3  # |g| h above this marks a surface particle (csf.py:53)
NORMAL_PASSES = 2  # extra Shepard passes of the colour gradient for the normal (csf.py:54)
CURVATURE_PASSES = 3  # Shepard passes of the curvature over the surface (csf.py:55)
XSPH_MIN = 0.3  # with viscosity "xsph", tension runs at xsph_c >= this (csf.py:56)
RH_RANGE_XSPH = (3.5, 7.0)  # R/h range of the laws, viscosity "xsph" (csf.py:57)
RH_RANGE_VISCOUS = (7.0, 10.0)  # R/h range of the laws, viscosity "weiler2018" (csf.py:58)
CSF_BAND = (18.0, 24.0)  # largest neighbour count nearby over which CSF takes over from the pair force
RAYLEIGH_FACTOR = 1.55  # the pair force's gamma calibration factor (tension.py:78)


def csf_weight(count, lo: float = CSF_BAND[0], hi: float = CSF_BAND[1]):
    """0 under ``lo`` neighbours, 1 from ``hi``, smoothstep between (csf_sparse.py:44-48)."""
    x = np.clip((np.asarray(count, np.float64) - lo) / (hi - lo), 0.0, 1.0)
    return x * x * (3.0 - 2.0 * x)


def cohesion_c_np(r, h: float) -> np.ndarray:
    """Akinci's cohesion spline C(r) = 32 / (pi h^9) {(h - r)^3 r^3 for h/2 < r <= h; 2 (h - r)^3 r^3 - h^6 / 64 for
    0 < r <= h/2} (tension.py:81-92, 172-181)."""
    r = np.asarray(r, dtype=np.float64)
    out = np.zeros_like(r)
    k = 32.0 / (math.pi * h**9)
    s = (h - r) ** 3 * r**3
    far = (r > 0.5 * h) & (r <= h)
    near = (r > 0.0) & (r <= 0.5 * h)
    out[far] = k * s[far]
    out[near] = k * (2.0 * s[near] - h**6 / 64.0)
    return out


def capillary_dt(spacing: float, rho0: float, sigma: float) -> float:
    """Capillary step limit sqrt(rho dx^3 / (2 pi sigma)) [s], inf without tension (tension.py:264-268; Brackbill,
    Kothe and Zemach 1992 with the particle spacing for the length)."""
    return math.inf if sigma <= 0.0 else math.sqrt(rho0 * spacing**3 / (2.0 * math.pi * sigma))


def _offsets(spacing: float, h: float) -> np.ndarray:
    n = int(math.ceil(h / spacing)) + 1
    ax = np.arange(-n, n + 1) * spacing
    return np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)


def flat_lattice_constants(spacing: float, h: float, kind: str = "cubic", layers: int = 6) -> tuple[float, float]:
    """(I, c_s) of a flat rest-lattice free surface (csf.py:252-278): I = the sum over the surface layers
    (|g| h > SURFACE_GRAD) of |g| spacing, the CSF delta normalisation; c_s = rho / rho0 extrapolated from the top
    layer half a spacing out to the interface."""
    off = _offsets(spacing, h)
    rr = np.linalg.norm(off, axis=1)
    keep = rr < h
    off, rr = off[keep], rr[keep]
    dl = np.round(off[:, 2] / spacing).astype(int)  # neighbour j sits on layer l - dl
    w = kernel_w_np(rr, h, kind)
    vol = 1.0 / float(w.sum())
    e = 1.0e-7 * h
    nz = rr > 0
    dw = np.zeros_like(rr)
    dw[nz] = (kernel_w_np(rr[nz] + e, h, kind) - kernel_w_np(rr[nz] - e, h, kind)) / (2.0 * e)
    gz_unit = np.where(nz, dw * (-off[:, 2]) / np.where(nz, rr, 1.0), 0.0)
    total, c0, g0 = 0.0, 0.0, 0.0
    for layer in range(layers):
        present = layer - dl >= 0
        gz = abs(float(np.sum(vol * gz_unit[present])))
        if gz * h > SURFACE_GRAD:
            total += gz * spacing
        if layer == 0:
            c0, g0 = float(np.sum(vol * w[present])), gz
    return total, c0 - 0.5 * spacing * g0


def _slab_pairs(spacing: float, h: float, kind: str, rho0: float, mass: float, layers: int):
    """Per top layer l < ``layers`` of the rest lattice's flat free surface: the offsets x_j - x_i to its neighbours
    and the per-unit-gamma pair forces on i (cohesion vector, curvature z) (tension.py:204-230)."""
    off = _offsets(spacing, h)
    rr = np.linalg.norm(off, axis=1)
    keep = (rr < h) & (rr > 0.0)
    off, rr = off[keep], rr[keep]
    dl = np.round(off[:, 2] / spacing).astype(int)
    depth = layers + 2 * int(math.ceil(h / spacing)) + 2
    w0 = float(kernel_w_np(np.zeros(1), h, kind)[0])
    rho = np.array([mass * (w0 + kernel_w_np(rr[lay - dl >= 0], h, kind).sum()) for lay in range(depth)])
    e = 1.0e-7 * h
    dw = (kernel_w_np(rr + e, h, kind) - kernel_w_np(rr - e, h, kind)) / (2.0 * e)
    grad_z = dw * (-off[:, 2]) / rr
    nz = np.zeros(depth)  # Akinci's normal (z) per layer: h sum_j m / rho_j dW
    for lay in range(depth - int(math.ceil(h / spacing)) - 1):
        ok = lay - dl >= 0
        nz[lay] = h * np.sum(mass / rho[lay - dl[ok]] * grad_z[ok])
    out = []
    for lay in range(layers):
        ok = lay - dl >= 0
        o, r, lj = off[ok], rr[ok], lay - dl[ok]
        K = 2.0 * rho0 / (rho[lay] + rho[lj])
        f_coh = (K * mass * mass * cohesion_c_np(r, h) / r)[:, None] * o
        f_curv_z = -K * mass * (nz[lay] - nz[lj]) * 4.0 * np.maximum(1.0 - r / h, 0.0)
        out.append((o, f_coh, f_curv_z))
    return out


def lattice_sigma_per_gamma(spacing: float, h: float, kind: str, rho0: float, mass: float, layers: int = 8) -> float:
    """Kirkwood-Buff surface tension [N/m] per unit gamma of the rest lattice's flat free surface (tension.py:241-249):
    sigma = (1/A) sum over the pairs across the surface of z_ij F_z - (x_ij F_x + y_ij F_y) / 2."""
    sigma = 0.0
    for o, f_coh, f_curv_z in _slab_pairs(spacing, h, kind, rho0, mass, layers):
        vir = -o[:, 2] * (f_coh[:, 2] + f_curv_z) - 0.5 * (-o[:, 0] * f_coh[:, 0] - o[:, 1] * f_coh[:, 1])
        sigma += 0.5 * float(vir.sum()) / spacing**2
    return sigma


def pair_gamma(sigma: np.ndarray, rho0: np.ndarray, mass: np.ndarray, spacing: float, h: float,
               kind: str) -> np.ndarray:
    """Per-particle gamma = sigma / (RAYLEIGH_FACTOR c_lattice(rho0, mass)), 0 where sigma is 0 (csf_sparse.py:91-99)."""
    sigma, rho0, mass = (np.asarray(a, np.float64) for a in (sigma, rho0, mass))
    out = np.zeros(max(len(sigma), 1), np.float32)
    for r0, m in {(float(a), float(b)) for a, b, s in zip(rho0, mass, sigma, strict=True) if s > 0.0}:
        sel = (rho0 == r0) & (mass == m) & (sigma > 0.0)
        out[:len(sigma)][sel] = sigma[sel] / (RAYLEIGH_FACTOR * lattice_sigma_per_gamma(spacing, h, kind, r0, m))
    return out


def shape_angles(contact_angle: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(cos, sin) of each shape's contact angle; sin = -1 marks a free contact line (angle -1) (csf.py:281-296)."""
    ang = np.asarray(contact_angle, np.float64)
    if len(ang) == 0:
        ang = np.full(1, -1.0)
    bad = (ang > math.pi + 1.0e-6) | ((ang < 0.0) & (ang != -1.0))
    if bad.any():
        raise ValueError(f"sph:shape_contact_angle must be -1 (free) or in [0, pi] rad; shapes {np.nonzero(bad)[0]}")
    free = ang < 0.0
    cos = np.where(free, 0.0, np.cos(ang)).astype(np.float32)
    sin = np.where(free, -1.0, np.sin(np.clip(ang, 0.0, math.pi))).astype(np.float32)
    return cos, sin


def xsph_coefficient(cfg, viscosity: str) -> float:
    """The XSPH coefficient the non-pressure update uses (solver.py:87 and 299): max(xsph_c, XSPH_MIN) with csf,
    and only with the viscosity "xsph"."""
    if viscosity != "xsph":
        return 0.0
    return max(cfg.xsph_c, XSPH_MIN) if cfg.surface_tension == "csf" else cfg.xsph_c


def liquid_gamma(model, cfg) -> np.ndarray:
    """``pair_gamma`` over the model's particles (0 off the SPH particles), as float32 [max(n, 1)]."""
    n = model.particle_count
    if n == 0:
        return np.zeros(1, np.float32)
    sph = model.sph.phase.cpu().numpy() >= 0
    sigma = np.where(sph, model.sph.surface_tension.cpu().numpy().astype(np.float64), 0.0)
    return pair_gamma(sigma, model.sph.rest_density.cpu().numpy(), model.particle_mass.cpu().numpy(), cfg.spacing,
                      cfg.h, cfg.kernel)


def capillary_limit(model, cfg) -> float:
    """The smallest capillary step over the liquid (csf.py:305-309)."""
    if model.particle_count == 0:
        return math.inf
    sph = model.sph.phase.cpu().numpy() >= 0
    rho0 = model.sph.rest_density.cpu().numpy()[sph]
    sigma = model.sph.surface_tension.cpu().numpy()[sph]
    return min((capillary_dt(cfg.spacing, float(r), float(s)) for r, s in zip(rho0, sigma, strict=True)),
               default=math.inf)


TENSION_INPUTS = {"none": (), "csf": ("surface_tension", "shape_contact_angle")}  # layout.py:74-77


def reject_unread_inputs(model, cfg) -> None:
    """Refuse tension inputs the configured model would silently drop (layout.py:80-95): nonzero
    sph:surface_tension / sph:adhesion on SPH particles, or a set sph:shape_contact_angle, that
    Config.surface_tension does not read."""
    used = TENSION_INPUTS.get(cfg.surface_tension, ())
    sph = model.sph.phase.cpu().numpy() >= 0
    for name, unset in (("surface_tension", 0.0), ("adhesion", 0.0), ("shape_contact_angle", -1.0)):
        arr = getattr(model.sph, name, None)
        per_particle = name != "shape_contact_angle"
        if name in used or arr is None or (per_particle and not np.any(sph)):
            continue
        vals = arr.cpu().numpy()[sph] if per_particle else np.asarray(arr)
        if np.any(vals != unset):
            what = "particles" if per_particle else "shapes"
            raise ValueError(f"SolverDFSPH: {np.count_nonzero(vals != unset)} {what} set sph:{name} but "
                             f"Config.surface_tension={cfg.surface_tension!r} does not read it "
                             f"(csf reads {TENSION_INPUTS['csf']})")
