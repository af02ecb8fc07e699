"""Continuum surface force on the GPU (csf.py, csf_sparse.py; tests/solvers/sph/test_csf.py and test_csf_sparse.py
of the reference): the HIP stage equal to the oracle bit for bit (free and sessile drops, contact angles, the viscous
solve after it, every tuning option), and the reference's laws on the HIP solver: Laplace pressure and the Rayleigh
l = 2 frequency within 10 %, the sessile contact angle within 5 degrees, free clusters gaining no energy."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.drops_util import (VISCOUS, WATER, cap, cap_fit_angle, capillary_step, drop_model, drop_points,
                              drop_radius, height_angle)

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

R = 1.0e-4


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sessile_model(radius_cells, liquid, contact_angle, start_angle=90.0, jitter=0.0, device="cuda"):
    sp = 2.0 * liquid.r
    rs, _ = cap(radius_cells * sp, contact_angle)
    base = rs * math.sin(math.radians(contact_angle)) if contact_angle < 90.0 else rs
    pts = drop_points(radius_cells, sp, hemisphere=True, cap_angle=start_angle)
    if jitter:
        pts = pts + np.random.default_rng(1).uniform(-jitter, jitter, pts.shape) * sp
    return drop_model(pts, liquid, plate_half=max(float(np.abs(pts[:, :2]).max()), base), contact_angle=contact_angle,
                      device=device)


def _damp(st, damping):
    st.particle_qd.mul_(damping)


@pytest.mark.parametrize("case", ["free", "sessile60", "viscous120"])
@pytest.mark.parametrize("options", [{}, {"graphs": 1, "check_interval": 2, "pair_cache": 7, "sparse_correction": 1}])
def test_drops_match_the_oracle_bitwise(case, options):
    """40 steps: positions, velocities, the surface flags, the applied curvature and the slot impulses (the contact
    line's wall reaction) equal the oracle's, bit for bit."""
    import torch

    from alloy_amd.sph import _abi
    from alloy_amd.sph.solver import SolverDFSPH
    from oracle.oracle import OracleDFSPH

    if case == "free":
        pts = drop_points(5.0, 2.0 * R) + np.random.default_rng(2).uniform(-0.02, 0.02, (1, 3)) * 2.0 * R
        m, cfg = drop_model(pts, device="cuda")
    elif case == "sessile60":
        m, cfg = _sessile_model(5.0, WATER, 60.0, jitter=0.02)
    else:
        m, cfg = _sessile_model(5.0, VISCOUS, 120.0, jitter=0.02)
    s, o = SolverDFSPH(m, cfg), OracleDFSPH(m, cfg, 32)
    for name, val in options.items():
        s.set_option(getattr(_abi, "OPT_" + name.upper()), val)
    st = m.state()
    n = m.particle_count
    dt = capillary_step(cfg, WATER)
    for k in range(40):
        s.step(st, st, None, None, dt)
        o.step(dt)
        torch.cuda.synchronize()
        what = f"{case} step {k}"
        gq, gv = st.particle_q.cpu().numpy(), st.particle_qd.cpu().numpy()
        assert np.array_equal(gq, o.q), f"{what}: positions differ at {np.nonzero((gq != o.q).any(1))[0][:8]}"
        assert np.array_equal(gv, o.qd), f"{what}: velocities differ at {np.nonzero((gv != o.qd).any(1))[0][:8]}"
        assert np.array_equal(s.buffer(_abi.BUF_CSF_SURF, np.int32, (n,)), o.buffer(_abi.BUF_CSF_SURF)[:n]), what
        assert np.array_equal(s.buffer(_abi.BUF_CSF_KAPPA, np.float32, (n,)), o.buffer(_abi.BUF_CSF_KAPPA)[:n]), what
        assert np.array_equal(s.buffer(_abi.BUF_SLOT_IMPULSE, np.float32, (12 * n,)),
                              o.buffer(_abi.BUF_SLOT_IMPULSE)[:12 * n]), what
    assert s.buffer(_abi.BUF_CSF_SURF, np.int32, (n,)).any()


def test_report_and_capillary_step():
    """test_csf.py:66-77: the report names csf, the step is capped at the capillary limit, a step above it is counted
    (never silent), and XSPH runs at least at XSPH_MIN."""
    from alloy_amd.sph import csf
    from alloy_amd.sph.solver import SolverDFSPH

    m, cfg = drop_model(drop_points(4.0, 2.0 * R), device="cuda")
    s = SolverDFSPH(m, cfg)
    assert s._params.xsph_c == pytest.approx(csf.XSPH_MIN)
    assert s.config.dt_max == pytest.approx(csf.capillary_dt(cfg.spacing, WATER.rho0, WATER.sigma))
    st = m.state()
    s.step(st, st, None, None, s.config.dt_max)
    rep = s.report()
    assert rep.tension == "csf" and not rep.tension_qualitative and rep.capillary_dt_violations == 0
    assert rep.tension_rh_range == csf.RH_RANGE_XSPH
    s.step(st, st, None, None, 2.0 * s.config.dt_max)
    assert s.report().capillary_dt_violations == 1


def test_low_viscosity_drops_warn(caplog):
    from alloy_amd.sph.solver import SolverDFSPH

    m, cfg = drop_model(drop_points(3.0, 2.0 * R), device="cuda")
    with caplog.at_level("WARNING", logger="alloy.sph"):
        SolverDFSPH(m, cfg)
    assert any("l = 2 oscillation grows" in r.message for r in caplog.records)


@pytest.mark.parametrize("radius_cells", [7.0, 10.0])
def test_laplace_pressure(radius_cells):
    """test_csf.py:96-103: the interior pressure of a resting water drop within 10 % of 2 sigma / R at R/h 3.5 and 5
    (600 damped settling steps, the mean over 150 more)."""
    import torch

    from alloy_amd.sph import _abi
    from alloy_amd.sph.solver import SolverDFSPH

    m, cfg = drop_model(drop_points(radius_cells, 2.0 * R), device="cuda")
    s = SolverDFSPH(m, cfg)
    n, dt = m.particle_count, s.config.dt_max
    big_r = drop_radius(n, float(m.particle_mass[0]), WATER.rho0)
    st = m.state()
    ps = []
    for k in range(750):
        s.step(st, st, None, None, dt)
        _damp(st, 0.9)
        if k >= 600:
            torch.cuda.synchronize()
            x = st.particle_q.cpu().numpy().astype(np.float64)
            inner = np.linalg.norm(x - x.mean(axis=0), axis=1) < big_r - 2.0 * cfg.h
            kk = s.buffer(_abi.BUF_KD, np.float32, (n,)) / dt**2 + s.buffer(_abi.BUF_KV, np.float32, (n,)) / dt
            ps.append(float(WATER.rho0 * kk[inner].mean()))
    ratio = float(np.mean(ps)) / (2.0 * WATER.sigma / big_r)
    _log(f"[csf] Laplace R/h {big_r / cfg.h:.2f}: p / (2 sigma / R) = {ratio:.4f}")
    assert abs(ratio - 1.0) <= 0.10, ratio


def _l2_moment(q):
    x = q - q.mean(axis=0)
    return float(np.mean(x[:, 2] ** 2) - 0.5 * np.mean(x[:, 0] ** 2 + x[:, 1] ** 2))


def _fit_damped(t, y, omega_guess):
    """(omega_d, lambda) of y = A exp(-lambda t) cos(omega_d t + phi) + c by least squares (drops.py:104-122)."""
    from scipy.optimize import curve_fit

    def f(t, A, lam, w, ph, c):
        return A * np.exp(-lam * t) * np.cos(w * t + ph) + c

    best = None
    for wg in (0.7 * omega_guess, omega_guess, 1.3 * omega_guess):
        try:
            p, _ = curve_fit(f, t, y, p0=[np.ptp(y) / 2, 50.0, wg, 0.0, float(np.mean(y))], maxfev=20000)
        except RuntimeError:
            continue
        res = float(np.sum((f(t, *p) - y) ** 2))
        if best is None or res < best[1]:
            best = (p, res)
    assert best is not None, "no fit converged"
    return abs(float(best[0][2])), float(best[0][1])


@pytest.mark.parametrize("radius_cells", [7.0, 10.0])
def test_rayleigh_l2_frequency(radius_cells):
    """test_csf.py:106-113 (drops.py:125-158): a settled water drop given the l = 2 velocity field; the undamped
    frequency sqrt(omega_d^2 + lambda^2) of the l = 2 moment within 10 % of Rayleigh's sqrt(8 sigma / (rho R^3))."""
    import torch

    from alloy_amd.sph.solver import SolverDFSPH

    m, cfg = drop_model(drop_points(radius_cells, 2.0 * R), device="cuda")
    s = SolverDFSPH(m, cfg)
    n, dt = m.particle_count, s.config.dt_max
    big_r = drop_radius(n, float(m.particle_mass[0]), WATER.rho0)
    w_th = math.sqrt(8.0 * WATER.sigma / (WATER.rho0 * big_r**3))
    st = m.state()
    for _ in range(600):
        s.step(st, st, None, None, dt)
        _damp(st, 0.9)
    torch.cuda.synchronize()
    q0 = st.particle_q.cpu().numpy().astype(np.float64)
    x = q0 - q0.mean(axis=0)
    v0 = 0.05 * w_th * big_r
    st.particle_qd.copy_(torch.tensor(v0 * np.stack([-x[:, 0], -x[:, 1], 2.0 * x[:, 2]], 1) / big_r,
                                      dtype=torch.float32))
    t, y = [], []
    period = 2.0 * math.pi / w_th
    for k in range(int(3.0 * period / dt) + 1):
        s.step(st, st, None, None, dt)
        if k % 2 == 0:
            t.append((k + 1) * dt)
            y.append(_l2_moment(st.particle_q.cpu().numpy().astype(np.float64)))
    w_d, lam = _fit_damped(np.array(t), np.array(y), w_th)
    ratio = math.hypot(w_d, lam) / w_th
    _log(f"[csf] Rayleigh R/h {big_r / cfg.h:.2f}: omega_0 / omega_2 = {ratio:.4f}")
    assert abs(ratio - 1.0) <= 0.10, ratio


@pytest.mark.parametrize("angle", [60.0, 120.0])
def test_sessile_contact_angle(angle):
    """test_csf.py:126-133 (drops.py:216-254): a hemisphere's volume of water at R/h 5 on a plate with contact angle
    ``angle``, lightly damped for 2000 steps: both measures (the cap height and the sphere fit) within 5 degrees."""
    import torch

    from alloy_amd.sph import _abi
    from alloy_amd.sph.solver import SolverDFSPH

    m, cfg = _sessile_model(10.0, WATER, angle)
    s = SolverDFSPH(m, cfg)
    n, dt = m.particle_count, s.config.dt_max
    volume = n * float(m.particle_mass[0]) / WATER.rho0
    st = m.state()
    for _ in range(2000):
        s.step(st, st, None, None, dt)
        _damp(st, 0.98)
    torch.cuda.synchronize()
    q = st.particle_q.cpu().numpy().astype(np.float64)
    surf = s.buffer(_abi.BUF_CSF_SURF, np.int32, (n,))
    theta, theta_fit = height_angle(q, volume, WATER.r), cap_fit_angle(q, surf, WATER.r, 2.0 * cfg.h)
    lost = int(np.sum(q[:, 2] < -WATER.r))
    _log(f"[csf] sessile {angle:.0f} deg: height {theta:.2f}, fit {theta_fit:.2f}, lost {lost}")
    assert lost == 0 and abs(theta - angle) <= 5.0 and abs(theta_fit - angle) <= 5.0


def test_a_free_cluster_at_rest_gains_no_energy_from_tension():
    """test_csf_sparse.py:34-39 for the 33-particle cluster: 1500 steps at the capillary step, the peak kinetic
    energy under 10 % of the cluster's surface energy."""
    import torch

    from alloy_amd.sph.solver import SolverDFSPH

    m, cfg = drop_model(drop_points(2.0, 2.0 * R), device="cuda")
    n, sp = m.particle_count, 2.0 * R
    assert n == 33
    radius = (3.0 * n * sp**3 / (4.0 * math.pi)) ** (1.0 / 3.0)
    e_surf = WATER.sigma * 4.0 * math.pi * radius**2 / (n * WATER.rho0 * sp**3)
    s = SolverDFSPH(m, cfg)
    st = m.state()
    peak = 0.0
    for k in range(1500):
        s.step(st, st, None, None, s.config.dt_max)
        if k % 10 == 9:
            torch.cuda.synchronize()
            peak = max(peak, 0.5 * float((st.particle_qd.cpu().numpy().astype(np.float64) ** 2).sum(axis=1).mean()))
    _log(f"[csf] 33-particle cluster: peak kinetic energy {peak / e_surf:.2%} of its surface energy")
    assert peak <= 0.10 * e_surf
