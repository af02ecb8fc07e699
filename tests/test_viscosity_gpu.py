"""Weiler 2018 implicit viscosity on the GPU (viscosity.py; tests/solvers/sph/test_viscosity.py of the reference): the
reference's physics checks (shear decay at nu k^2, momentum under rigid motion, a viscous column at rest) on the HIP
solver, and the HIP step equal to the oracle bit for bit, per world, under every tuning option."""

from __future__ import annotations

import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

MU, DT = 1.0, 1.25e-3


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _blob(counts, viscosity, device="cuda"):
    from alloy_amd.sph.build import add_fluid_block
    from alloy_amd.sph.config import DFSPHConfig
    from alloy_amd.sph.model import ModelBuilder

    cfg = DFSPHConfig(particle_radius=1.0e-3, viscosity=viscosity)
    cfg.validate()
    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    b.begin_world()
    add_fluid_block(b, cfg, (0.0, 0.0, 0.0), counts, viscosity=MU)
    b.end_world()
    return b.finalize(device), cfg


def _columns(mus, fills, nx=4, ny=4, headroom=6):
    """One resting column per world in a box tank (walls h thick), each world with its own viscosity."""
    from alloy_amd.sph import scenes
    from alloy_amd.sph.build import add_box_tank, add_fluid_block
    from alloy_amd.sph.config import DFSPHConfig
    from alloy_amd.sph.model import ModelBuilder

    cfg = DFSPHConfig(particle_radius=1.0e-3, viscosity="weiler2018")
    cfg.validate()
    s = cfg.spacing
    lo, hi = (0.0, 0.0, 0.0), (nx * s, ny * s, (max(fills) + headroom) * s)
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    for mu, f in zip(mus, fills):
        b.begin_world()
        add_box_tank(b, lo, hi, wall=cfg.h)
        add_fluid_block(b, cfg, lo, (nx, ny, f), viscosity=mu)
        b.end_world()
    return scenes.Scene(b.finalize("cuda"), cfg)


def test_shear_mode_decays_at_nu_k2():
    import torch

    from alloy_amd.sph.solver import SolverDFSPH

    out = {}
    for visc in ("weiler2018", "xsph"):
        m, cfg = _blob((12, 12, 24), visc)
        s = SolverDFSPH(m, cfg)
        st = m.state()
        q = st.particle_q.cpu().numpy()
        k = 2.0 * np.pi / (24 * cfg.spacing)
        st.particle_qd[:, 0] = torch.tensor(0.01 * np.sin(k * q[:, 2]), dtype=torch.float32, device="cuda")
        hi = np.array([12, 12, 24]) * cfg.spacing
        inner = np.all((q > 2 * cfg.h) & (q < hi - 2 * cfg.h), axis=1)

        def amp():
            torch.cuda.synchronize()
            qq, vv = st.particle_q.cpu().numpy(), st.particle_qd.cpu().numpy()
            return np.linalg.lstsq(np.sin(k * qq[inner, 2])[:, None], vv[inner, 0], rcond=None)[0][0]

        a0 = amp()
        for _ in range(10):
            s.step(st, st, None, None, DT)
        out[visc] = amp() / a0
    ratio = np.log(out["weiler2018"]) / (-10.0 * np.log1p(MU / 1000.0 * k * k * DT))
    _log(f"[viscosity] shear decay exponent / analytic = {ratio:.4f}, xsph keeps {out['xsph']:.5f}")
    assert 0.7 < ratio < 1.0
    assert out["xsph"] > 0.99


def test_rigid_motion_conserves_momentum():
    import torch

    from alloy_amd.sph.solver import SolverDFSPH

    m, cfg = _blob((10, 10, 10), "weiler2018")
    s = SolverDFSPH(m, cfg)
    st = m.state()
    q = st.particle_q.cpu().numpy().astype(np.float64)
    c = q.mean(axis=0)
    v = np.cross(np.array([0.0, 0.0, 2.0]), q - c) + np.array([0.01, 0.0, 0.0])
    st.particle_qd.copy_(torch.tensor(v, dtype=torch.float32))
    v = st.particle_qd.cpu().numpy().astype(np.float64)
    mass = m.particle_mass.cpu().numpy().astype(np.float64)[:, None]
    p0, l0 = (mass * v).sum(0), (mass * np.cross(q - c, v)).sum(0)
    for _ in range(10):
        s.step(st, st, None, None, DT)
    torch.cuda.synchronize()
    q1, v1 = st.particle_q.cpu().numpy().astype(np.float64), st.particle_qd.cpu().numpy().astype(np.float64)
    p1, l1 = (mass * v1).sum(0), (mass * np.cross(q1 - q1.mean(0), v1)).sum(0)
    assert np.allclose(p1, p0, rtol=1e-4, atol=1e-10)
    assert l1[2] == pytest.approx(l0[2], rel=1e-4)


def test_viscous_column_rests():
    """test_viscosity.py:93-101: a 6 x 6 x 16 column at 1 Pa s, primed, 100 steps: nothing escapes and the fastest
    particle stays under 5 cm/s."""
    from alloy_amd.sph import scenes
    from alloy_amd.sph.config import DFSPHConfig

    sc = scenes.water_column(6, 6, 16, cfg=DFSPHConfig(particle_radius=1.0e-3, viscosity="weiler2018"),
                             device="cuda")
    sc.model.sph.viscosity.fill_(MU)
    s = sc.make_solver()
    assert s.viscosity == "weiler2018"
    st = sc.model.state()
    s.prime(st, DT)
    for _ in range(100):
        s.step(st, st, None, None, DT)
    r = s.report()
    _log(f"[viscosity] column max speed {r.max_speed[0]:.4g} m/s")
    assert r.escaped[0] == 0 and r.max_speed[0] < 0.05


def _compare(s, o, st, what):
    import torch

    from alloy_amd.sph import _abi

    torch.cuda.synchronize()
    W = s.model.world_count
    gq, gv = st.particle_q.cpu().numpy(), st.particle_qd.cpu().numpy()
    assert np.array_equal(gq, o.q), f"{what}: positions differ at {np.nonzero((gq != o.q).any(1))[0][:8]}"
    assert np.array_equal(gv, o.qd), f"{what}: velocities differ at {np.nonzero((gv != o.qd).any(1))[0][:8]}"
    for which in (_abi.BUF_VISC_ITERS, _abi.BUF_ITERS_DEN, _abi.BUF_ITERS_DIV):
        assert np.array_equal(s.buffer(which, np.int32, (W,)), o.buffer(which)[:W]), f"{what}: buffer {which}"


def test_blob_matches_the_oracle_bitwise():
    import torch

    from alloy_amd.sph import _abi
    from alloy_amd.sph.solver import SolverDFSPH
    from oracle.oracle import OracleDFSPH

    m, cfg = _blob((8, 8, 8), "weiler2018")
    rng = np.random.default_rng(11)
    v = rng.normal(0.0, 0.02, (m.particle_count, 3)).astype(np.float32)
    s, o = SolverDFSPH(m, cfg), OracleDFSPH(m, cfg, 32)
    st = m.state()
    st.particle_qd.copy_(torch.tensor(v))
    o.qd[...] = v
    iters = []
    for k in range(10):
        s.step(st, st, None, None, DT)
        o.step(DT)
        _compare(s, o, st, f"blob step {k}")
        iters.append(int(o.buffer(_abi.BUF_VISC_ITERS)[0]))
    _log(f"[viscosity] blob viscous iterations {iters}")
    assert min(iters) > 0


@pytest.mark.parametrize("options", [{}, {"graphs": 1, "check_interval": 3, "pair_cache": 5, "sparse_correction": 1}])
def test_columns_per_world_viscosity_match_the_oracle_bitwise(options):
    """Two worlds at 1 and 0.05 Pa s (different heights) in SDF tanks, primed, 20 steps: the per-world solves stop
    on their own iteration counts and every step equals the oracle's."""
    from alloy_amd.sph import _abi
    from alloy_amd.sph.solver import hydrostatic_guess
    from oracle.oracle import OracleDFSPH

    sc = _columns([1.0, 0.05], [8, 6])
    s = sc.make_solver()
    for name, val in options.items():
        s.set_option(getattr(_abi, "OPT_" + name.upper()), val)
    o = OracleDFSPH(sc.model, sc.config, 32)
    st = sc.model.state()
    s.prime(st, DT)
    o.prime(hydrostatic_guess(sc.model, sc.model.state(), sc.config.spacing), DT)
    _compare(s, o, st, "prime")
    iters = []
    for k in range(20):
        s.step(st, st, None, None, DT)
        o.step(DT)
        _compare(s, o, st, f"step {k}")
        iters.append(o.buffer(_abi.BUF_VISC_ITERS)[:2].tolist())
    _log(f"[viscosity] per-world viscous iterations {iters[-3:]}")
    assert any(a != b for a, b in iters)


# ---------------------------------------------------------------------------------------------- the wall term
def _film_gpu(e, mu_b, T, options=None):
    import torch

    from alloy_amd.sph import _abi
    from alloy_amd.sph.solver import SolverDFSPH
    from tests.test_oracle_wall_viscosity import film_measure, film_model

    m, cfg = film_model(e, mu_b, device="cuda")
    s = SolverDFSPH(m, cfg)
    for name, val in (options or {}).items():
        s.set_option(getattr(_abi, "OPT_" + name.upper()), val)
    st = m.state()
    dt = 2.0 * cfg.particle_radius
    for _ in range(int(round(T / dt))):
        s.step(st, st, None, None, dt)
    torch.cuda.synchronize()
    return film_measure(st.particle_q.cpu().numpy().astype(np.float64),
                        st.particle_qd.cpu().numpy().astype(np.float64), cfg)


@pytest.mark.parametrize("e", [4.0e-3, 2.0e-3, 1.0e-3])
def test_film_drains_at_the_nusselt_speed(e):
    """test_wall_viscosity.py:164-171: honey films 1, 2 and 4 mm thick on a static wall drain within 15 % of the
    Nusselt speed rho g e^2 / (3 mu); nothing enters the wall."""
    from tests.test_oracle_wall_viscosity import G, HONEY_MU, HONEY_RHO

    f = _film_gpu(e, HONEY_MU, 0.05)
    nusselt = HONEY_RHO * G * f["e"] ** 2 / (3.0 * HONEY_MU)
    _log(f"[wall viscosity] e = {e * 1e3:.0f} mm: u / Nusselt = {f['u'] / nusselt:.4f}")
    assert abs(f["u"] - nusselt) <= 0.15 * nusselt, (f, nusselt)
    assert f["x_min"] > 0.0, f


def test_ten_times_wall_viscosity_stays_stable():
    """test_wall_viscosity.py:185-194: 10x mu_b on a 2 mm film for 0.6 s stays out of the wall with bounded speeds,
    never faster than no-slip drainage."""
    from tests.test_oracle_wall_viscosity import G, HONEY_MU, HONEY_RHO

    f = _film_gpu(2.0e-3, 10.0 * HONEY_MU, 0.6)
    nusselt = HONEY_RHO * G * f["e"] ** 2 / (3.0 * HONEY_MU)
    _log(f"[wall viscosity] 10x mu_b: u / Nusselt = {f['u'] / nusselt:.4f}, v_max / Nusselt = {f['v_max'] / nusselt:.3f}")
    assert f["x_min"] > 0.0, f
    assert f["v_max"] < 3.0 * nusselt, (f, nusselt)
    assert 0.0 < f["u"] <= 1.15 * nusselt, (f, nusselt)


def test_boundary_viscosity_needs_the_implicit_viscosity():
    from alloy_amd.sph.solver import SolverDFSPH
    from tests.test_oracle_wall_viscosity import film_model

    m, cfg = film_model(4.0e-3, 6.0, ny=4, nz=4, device="cuda")
    import dataclasses

    m.sph.viscosity.fill_(1.0e-3)  # 'auto' resolves to xsph below 1e-2 Pa s
    with pytest.raises(ValueError, match="boundary_viscosity needs the weiler2018"):
        SolverDFSPH(m, dataclasses.replace(cfg, viscosity="auto"))


@pytest.mark.parametrize("options", [{}, {"graphs": 1, "check_interval": 2, "pair_cache": 7, "sparse_correction": 1}])
def test_film_matches_the_oracle_bitwise(options):
    """A 2 mm honey film with the wall term, 30 steps (the pressure guess shifts the wall velocity from the second
    step on): positions, velocities, iteration counts, the slot impulses and the shape meters equal the oracle's."""
    import torch

    from alloy_amd.sph import _abi
    from alloy_amd.sph.solver import SolverDFSPH
    from oracle.oracle import OracleDFSPH
    from tests.test_oracle_wall_viscosity import film_model

    m, cfg = film_model(2.0e-3, 6.0, ny=8, nz=24, device="cuda")
    s, o = SolverDFSPH(m, cfg), OracleDFSPH(m, cfg, 32)
    for name, val in options.items():
        s.set_option(getattr(_abi, "OPT_" + name.upper()), val)
    st = m.state()
    n = m.particle_count
    dt = 2.0 * cfg.particle_radius
    for k in range(30):
        s.step(st, st, None, None, dt)
        o.step(dt)
        _compare(s, o, st, f"film step {k}")
        assert np.array_equal(s.buffer(_abi.BUF_SLOT_IMPULSE, np.float32, (n * 12,)),
                              o.buffer(_abi.BUF_SLOT_IMPULSE)[:n * 12]), f"film step {k}: slot impulses"
    torch.cuda.synchronize()
    S = m.shape_count
    assert np.array_equal(s.buffer(_abi.BUF_SHAPE_METER, np.int64, (S * 3,)), o.buffer(_abi.BUF_SHAPE_METER)[:S * 3])


def test_primed_columns_with_wall_viscosity_match_the_oracle_bitwise():
    """Two worlds with the wall term (1 and 0.05 Pa s, mu_b = 1) in SDF tanks, primed (no pressure guess), 20 steps."""
    import dataclasses

    from alloy_amd.sph import _abi
    from alloy_amd.sph.solver import hydrostatic_guess
    from oracle.oracle import OracleDFSPH

    sc = _columns([1.0, 0.05], [8, 6])
    sc.config = dataclasses.replace(sc.config, boundary_viscosity=1.0)
    s = sc.make_solver()
    o = OracleDFSPH(sc.model, sc.config, 32)
    st = sc.model.state()
    s.prime(st, DT)
    o.prime(hydrostatic_guess(sc.model, sc.model.state(), sc.config.spacing), DT)
    _compare(s, o, st, "prime")
    for k in range(20):
        s.step(st, st, None, None, DT)
        o.step(DT)
        _compare(s, o, st, f"step {k}")
    assert s.report().max_speed.max() < 0.05


def test_pour_worlds_with_wall_viscosity_match_the_oracle_bitwise():
    """Four pour worlds (spout, swaying particle-wall bowl) at 1 Pa s with the wall term, 80 steps: the moving bowl's
    slot velocities enter the wall term; every 20 steps the GPU state equals the oracle's."""
    import dataclasses

    import torch

    from alloy_amd.sph import scenes
    from tests.test_emitters_gpu import _oracle_for

    base = scenes.pour_worlds(2048, seed=4, device="cuda", subset=range(0, 2048, 512))
    cfg = dataclasses.replace(base.config, viscosity="weiler2018", boundary_viscosity=1.0)
    sc = dataclasses.replace(base, config=cfg)
    sc.model.sph.viscosity.fill_(1.0)
    s = sc.make_solver()
    st = sc.model.state()
    o = _oracle_for(s, sc)
    for k in range(80):
        sc.drive(st, k * DT)
        o.qd[...] = st.particle_qd.cpu().numpy()
        s.step(st, st, None, None, DT)
        o.step(DT)
        if k % 20 == 19:
            _compare(s, o, st, f"pour step {k + 1}")
    torch.cuda.synchronize()
