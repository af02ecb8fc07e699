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

from .kernels import kernel_dwdr_np, kernel_w_np

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


class _Planes:
    """The rest lattice's neighbours within h of a site, grouped by lattice plane: plane dz (from -m to m, m =
    ceil(h / spacing)) holds the offsets o = x_j - x_i with o_z = dz spacing. A flat free surface with the liquid on
    layers l = 0, 1, 2, ... (layer l at z = -l spacing) gives a site of layer l exactly the planes dz <= l, so every
    per-layer sum of the surface constants is a running sum of per-plane sums over dz.

    Per plane: ``w`` = sum W(|o|) (the site itself included in plane 0), ``w_pair`` the same without the site,
    ``gz`` = sum over o != 0 of the z component of grad_i W(x_i - x_j) = W'(r) (-o_z) / r, ``coh`` = sum C(r) / r
    (|o_xy|^2 / 2 - o_z^2) (the cohesion spline's virial weight), ``taper`` = sum 4 max(1 - r / h, 0) (the
    curvature term's weight)."""

    def __init__(self, spacing: float, h: float, kind: str):
        self.m = int(math.ceil(h / spacing))
        self.dz = np.arange(-self.m, self.m + 1)
        k = np.arange(-self.m - 1, self.m + 2, dtype=np.float64) * spacing
        ox, oy = np.meshgrid(k, k, indexing="ij")
        ox, oy = ox.ravel(), oy.ravel()
        self.w, self.w_pair, self.gz, self.coh, self.taper = (np.zeros(len(self.dz)) for _ in range(5))
        for p, d in enumerate(self.dz):
            oz = d * spacing
            r = np.sqrt(ox * ox + oy * oy + oz * oz)
            inside = r < h
            pair = inside & (r > 0.0)
            self.w[p] = kernel_w_np(r[inside], h, kind).sum()
            self.w_pair[p] = kernel_w_np(r[pair], h, kind).sum()
            rp, xy2 = r[pair], ox[pair] ** 2 + oy[pair] ** 2
            self.gz[p] = np.sum(kernel_dwdr_np(rp, h, kind) * (-oz) / rp)
            self.coh[p] = np.sum(cohesion_c_np(rp, h) / rp * (0.5 * xy2 - oz * oz))
            self.taper[p] = np.sum(4.0 * np.maximum(1.0 - rp / h, 0.0))

    def upto(self, values: np.ndarray, layer: int) -> float:
        """sum of the per-plane ``values`` over the planes a site of ``layer`` sees (dz <= layer)."""
        return float(values[self.dz <= layer].sum())


def flat_lattice_constants(spacing: float, h: float, kind: str = "cubic", layers: int = 6) -> tuple[float, float]:
    """(I, c_s) of a flat rest-lattice free surface (csf.py:225-250). With the volume V = 1 / (lattice kernel sum),
    layer l has the colour c_l = V sum_{dz <= l} w and the colour gradient g_l = V sum_{dz <= l} gz (z only, by
    symmetry). I, the CSF delta normalisation, is the gradient integrated across the surface layers: spacing times
    the sum of |g_l| over the layers with |g_l| h > SURFACE_GRAD. c_s is rho / rho0 at the interface, c_0
    extrapolated half a spacing outward along the gradient: c_0 - spacing |g_0| / 2."""
    pl = _Planes(spacing, h, kind)
    vol = 1.0 / float(pl.w.sum())
    g = np.array([abs(vol * pl.upto(pl.gz, lay)) for lay in range(layers)])
    big_i = spacing * float(g[g * h > SURFACE_GRAD].sum())
    return big_i, vol * pl.upto(pl.w, 0) - 0.5 * spacing * float(g[0])


def lattice_sigma_per_gamma(spacing: float, h: float, kind: str, rho0: float, mass: float, layers: int = 8) -> float:
    """Kirkwood-Buff surface tension [N/m] per unit gamma of the rest lattice's flat free surface (tension.py:239-249).

    The pair force on i from j (per unit gamma, csf_sparse.py's small-piece force) is K_ij (m^2 C(r) o / r - m
    (n_i - n_j) 4 max(1 - r/h, 0)) with o = x_j - x_i, K_ij = 2 rho0 / (rho_i + rho_j), the layer densities rho_l =
    m (W(0) + sum_{dz <= l} w_pair) and Akinci's normals n_l = h m sum_{dz <= l} gz / rho_{l - dz} (z only). The
    virial sigma = 1 / (2 A) sum_i sum_j [z_ij F_z - (x_ij F_x + y_ij F_y) / 2], A = spacing^2 per lattice column,
    x_ij = -o, then reduces per plane to the ``coh`` and ``taper`` sums. Pairs deeper than h below the surface cancel
    by the lattice's cubic symmetry, so ``layers`` > h / spacing + 2 covers every term."""
    pl = _Planes(spacing, h, kind)
    depth = layers + 2 * pl.m + 2
    w0 = float(kernel_w_np(np.zeros(1), h, kind)[0])
    rho = np.array([mass * (w0 + pl.upto(pl.w_pair, lay)) for lay in range(depth)])
    normal = np.zeros(depth)
    for lay in range(depth - pl.m - 1):  # the layers whose every neighbour plane has a density
        seen = pl.dz <= lay
        normal[lay] = h * mass * float(np.sum(pl.gz[seen] / rho[lay - pl.dz[seen]]))
    sigma = 0.0
    for lay in range(layers):
        for p in np.nonzero(pl.dz <= lay)[0]:
            d = int(pl.dz[p])
            k_ij = 2.0 * rho0 / (rho[lay] + rho[lay - d])
            curv = mass * (normal[lay] - normal[lay - d]) * d * spacing * pl.taper[p]
            sigma += k_ij * (mass * mass * pl.coh[p] + curv)
    return float(0.5 * sigma / spacing**2)


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
    """(cos, sin) of each shape's contact angle as float32, for the Breinlinger turn of the normal at a contact line
    (csf.py:253-266). An angle of exactly -1 marks a free contact line, stored as sin = -1 (a sine no angle in
    [0, pi] has); any other negative angle or one beyond pi is refused. No shapes: one free entry."""
    theta = np.asarray(contact_angle, np.float64).reshape(-1)
    if theta.size == 0:
        theta = np.array([-1.0])
    free = theta == -1.0
    valid = free | ((theta >= 0.0) & (theta <= math.pi + 1.0e-6))
    if not valid.all():
        raise ValueError(f"sph:shape_contact_angle takes -1 (a free contact line) or an angle in [0, pi] rad; "
                         f"shapes {np.flatnonzero(~valid).tolist()} are outside")
    turned = np.clip(theta, 0.0, math.pi)
    return (np.where(free, 0.0, np.cos(turned)).astype(np.float32),
            np.where(free, -1.0, np.sin(turned)).astype(np.float32))


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
