"""The oracle's smoothing kernels and wall volume against closed forms and independent quadratures
(tests/solvers/sph/test_kernels.py of the reference holds the same properties), and the host constants against the
reference's A14 pack (tests/golden/a14_constants.json)."""

from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "a14_constants.json")
KINDS = [0, 1]


def _radial_integral(f, h, n=20000):
    """int_0^h f(r) dr by the midpoint rule."""
    r = (np.arange(n) + 0.5) * (h / n)
    return float(sum(f(x) for x in r) * (h / n))


@pytest.mark.parametrize("kind", KINDS)
def test_kernel_integrates_to_one(kind):
    h = 1.0
    total = _radial_integral(lambda r: 4.0 * math.pi * r * r * orc.kernel_w(r, h, kind), h)
    assert abs(total - 1.0) < 1e-6


@pytest.mark.parametrize("kind", KINDS)
def test_kernel_support_and_values(kind):
    h = 0.02
    assert orc.kernel_w(h, h, kind) == 0.0 and orc.kernel_w(1.5 * h, h, kind) == 0.0
    w0 = orc.kernel_w(0.0, h, kind)
    expect = 8.0 / (math.pi * h**3) if kind == 0 else 21.0 / (2.0 * math.pi * h**3)
    assert w0 == pytest.approx(expect, rel=1e-14)
    # cubic: continuous at q = 1/2 (both pieces give k/4)
    if kind == 0:
        assert orc.kernel_w(0.5 * h, h, 0) == pytest.approx(expect / 4.0, rel=1e-12)


@pytest.mark.parametrize("kind", KINDS)
def test_kernel_gradient_matches_finite_difference(kind, rng):
    h = 1.0
    e = 1.0e-6
    for x in rng.uniform(-h, h, (200, 3)):
        g = orc.kernel_grad(x, h, kind)
        fd = np.zeros(3)
        for a in range(3):
            dx = np.zeros(3)
            dx[a] = e
            fd[a] = (orc.kernel_w(np.linalg.norm(x + dx), h, kind) - orc.kernel_w(np.linalg.norm(x - dx), h, kind)) / (2 * e)
        assert np.linalg.norm(g - fd) <= 1e-6 * max(np.linalg.norm(fd), 1.0)
    assert np.all(orc.kernel_grad([0.0, 0.0, 0.0], h, kind) == 0.0)


def _psi_quadrature(d: float, h: float, kind: int) -> float:
    """Kernel mass in the solid half-space z < -d: sphere shells of radius r contribute the surface fraction below
    the plane, (1 - d / r) / 2 for |d| < r."""
    def frac(r):
        if r <= d:
            return 0.0
        if r <= -d:
            return 1.0
        return 0.5 * (1.0 - d / r)
    return _radial_integral(lambda r: 4.0 * math.pi * r * r * orc.kernel_w(r, h, kind) * frac(r), h, 8000)


@pytest.mark.parametrize("kind", KINDS)
def test_halfspace_volume_matches_quadrature(kind):
    h = 1.0
    for d in (-1.5, -0.6, -0.25, 0.0, 0.25, 0.5, 0.8, 1.2):
        psi, _ = orc.halfspace_volume(d, h, 0.0, kind)
        assert abs(psi - _psi_quadrature(d, h, kind)) < 2e-6, d
    assert orc.halfspace_volume(-1.5, h, 0.0, kind)[0] == pytest.approx(1.0, abs=1e-12)
    assert orc.halfspace_volume(0.0, h, 0.0, kind)[0] == pytest.approx(0.5, abs=1e-12)
    assert orc.halfspace_volume(1.2, h, 0.0, kind) == (0.0, 0.0)


@pytest.mark.parametrize("kind", KINDS)
def test_halfspace_slope_matches_finite_difference(kind):
    h, e = 1.0, 1e-6
    for d in np.linspace(-0.95, 0.95, 39):
        _, dpsi = orc.halfspace_volume(d, h, 0.3, kind)
        fd = (orc.halfspace_volume(d + e, h, 0.3, kind)[0] - orc.halfspace_volume(d - e, h, 0.3, kind)[0]) / (2 * e)
        assert dpsi == pytest.approx(fd, abs=1e-6)


def test_curvature_lowers_convex_volume():
    flat = orc.halfspace_volume(0.3, 1.0, 0.0, 0)[0]
    convex = orc.halfspace_volume(0.3, 1.0, -1.0, 0)[0]
    assert convex < flat


def test_float32_kernels_track_float64():
    h = 0.02
    for r in np.linspace(0.0, h, 50):
        assert orc.kernel_w(r, h, 0, 32) == pytest.approx(orc.kernel_w(r, h, 0, 64), rel=2e-6, abs=1e-3)


def _lattice(spacing, h):
    m = int(math.ceil(h / spacing)) + 1
    ax = np.arange(-m, m + 1) * spacing
    return np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)


def test_host_constants_match_reference_pack():
    """S, m = rho0/S and the lattice neighbour count at r = 1 mm, as the reference's A14 pack states them."""
    from alloy_amd.sph import kernels as kn
    from alloy_amd.sph.config import DFSPHConfig, full_neighbor_count, particle_mass

    with open(GOLDEN) as f:
        g = json.load(f)
    cfg = DFSPHConfig(particle_radius=g["particle_radius"])
    S = kn.lattice_sum(cfg.spacing, cfg.h)
    assert S == pytest.approx(g["lattice_sum_S"], rel=1e-8)
    assert particle_mass(cfg, g["rest_density"]) == pytest.approx(g["particle_mass"], rel=1e-12)
    assert full_neighbor_count(cfg) == g["lattice_neighbours"]
    # the same sum through the oracle's own kernel
    pts = _lattice(cfg.spacing, cfg.h)
    S_orc = sum(orc.kernel_w(float(np.linalg.norm(p)), cfg.h, 0) for p in pts)
    assert S_orc == pytest.approx(S, rel=1e-12)
    # the pack's initial lattice sits at cell centres (i + 1/2) spacing
    q = np.asarray(g["first_positions"])
    frac = q / cfg.spacing - 0.5
    assert np.allclose(frac, np.round(frac), atol=1e-4)


def test_boundary_fit_reproduces_half_lattice_deficit():
    """A Psi(r + delta) equals the deficit of a lattice beside a flat wall (rest density exact at walls), with the
    deficit computed by a direct lattice sum of the oracle's kernel."""
    from alloy_amd.sph import kernels as kn

    sp, h = 2.0e-3, 4.0e-3
    delta, amp = kn.boundary_fit(sp, h)
    pts = _lattice(sp, h)
    w = np.array([orc.kernel_w(float(np.linalg.norm(p)), h, 0) for p in pts])
    deficit = 1.0 - w[pts[:, 2] >= 0].sum() / w.sum()
    psi, _ = orc.halfspace_volume(0.5 * sp + delta, h, 0.0, 0)
    assert amp * psi == pytest.approx(deficit, rel=1e-6)


def test_lattice_diagonal_matches_oracle_gradients():
    from alloy_amd.sph import kernels as kn

    sp, h = 0.01, 0.02
    v = 1.0 / kn.lattice_sum(sp, h)
    diag = sum(float(np.dot(g, g)) for g in (v * orc.kernel_grad(p, h, 0) for p in _lattice(sp, h)))
    assert kn.lattice_diagonal(sp, h) == pytest.approx(diag, rel=1e-6)


def test_box_sdf_special_cases():
    half = [0.1, 0.2, 0.3]
    phi, g = orc.shape_sdf(7, half, [0.15, 0.0, 0.0])
    assert phi == pytest.approx(0.05) and np.allclose(g, [1, 0, 0])
    phi, g = orc.shape_sdf(7, half, [0.0, 0.0, -0.25])
    assert phi == pytest.approx(-0.05) and np.allclose(g, [0, 0, -1])
    phi, g = orc.shape_sdf(7, half, [0.13, 0.24, 0.0])  # outside an edge
    assert phi == pytest.approx(0.05) and np.allclose(g, [0.6, 0.8, 0.0])
    phi, g = orc.shape_sdf(3, [0.5, 0, 0], [0.0, 0.0, 0.7])
    assert phi == pytest.approx(0.2) and np.allclose(g, [0, 0, 1])
