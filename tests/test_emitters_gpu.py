"""Emitters, sinks, the escape box and the int64 ledger on the GPU (emitters.py:67-172, runtime.py:197-347 of the
reference), bit for bit against the oracle: positions, velocities, flags, phase, the ledgers, the emitters' clocks and
cursors, over whole steps. Then the reference's own checks through the plug-in surface (test_mass.py,
test_escape_box.py, test_flux_energy.py) and a pack of pour worlds with swaying particle-wall bowls."""

from __future__ import annotations

import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DT = 1.0e-3


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _exact(led):
    return np.array_equal(led["active"] + led["out"], led["initial"] + led["in"])


def _oracle_for(solver, sc):
    from oracle.oracle import OracleDFSPH

    o = OracleDFSPH(sc.model, sc.config, 32)
    solver.emitters.flush()  # the solver uploads its table lazily (at the next step); the oracle needs it now
    if solver.emitters.struct is not None:
        o.set_emitters(solver.emitters.struct)
    return o


def _compare(solver, o, st, what):
    import torch

    from alloy_amd.sph import _abi

    torch.cuda.synchronize()
    n, W = solver.model.particle_count, solver.model.world_count
    gq, gv = st.particle_q.cpu().numpy(), st.particle_qd.cpu().numpy()
    assert np.array_equal(gq, o.q), f"{what}: positions differ at {np.nonzero((gq != o.q).any(1))[0][:8]}"
    assert np.array_equal(gv, o.qd), f"{what}: velocities differ at {np.nonzero((gv != o.qd).any(1))[0][:8]}"
    assert np.array_equal(solver.model.particle_flags.cpu().numpy(), o.buffer(_abi.BUF_FLAGS)[:n]), what
    assert np.array_equal(solver.model.sph.phase.cpu().numpy(), o.buffer(_abi.BUF_PHASE)[:n]), what
    for which, dt in ((_abi.BUF_MASS_IN, np.int64), (_abi.BUF_MASS_OUT, np.int64), (_abi.BUF_ESCAPED, np.int32),
                      (_abi.BUF_EMIT_OVERFLOW, np.int32), (_abi.BUF_FLUX_E, np.int64), (_abi.BUF_FLUX_G, np.int64),
                      (_abi.BUF_MAX_SPEED, np.float32), (_abi.BUF_ITERS_DEN, np.int32)):
        assert np.array_equal(solver.buffer(which, dt, (W,)), o.buffer(which)[:W]), f"{what}: buffer {which}"
    E = len(solver.emitters.table.world)
    if E:
        assert np.array_equal(solver.buffer(_abi.BUF_EMIT_CURSOR, np.int32, (E,)), o.buffer(_abi.BUF_EMIT_CURSOR))
        assert np.array_equal(solver.buffer(_abi.BUF_EMIT_T, np.float32, (E,)), o.buffer(_abi.BUF_EMIT_T))
    if solver.pw.count:  # the particle walls' share of the slot impulses and their work (particle_walls.py:138-177)
        assert np.array_equal(solver.buffer(_abi.BUF_PW_EXCH, np.int64, (solver.pw.count * 3,)),
                              o.buffer(_abi.BUF_PW_EXCH)), f"{what}: wall impulses"
        assert np.array_equal(solver.buffer(_abi.BUF_PW_WORK, np.int64, (W,)), o.buffer(_abi.BUF_PW_WORK)[:W])
    gl, ol = solver.mass_ledger(), o.mass_ledger()
    for k in ("initial", "in", "out", "active"):
        assert np.array_equal(gl[k], ol[k]), f"{what}: ledger {k}"
    assert _exact(gl)


def test_pour_with_sink_and_escape_box_matches_the_oracle_bitwise():
    """LIQ-1's scene (test_mass.py:13-41): an emitter pouring into world 0, a sink in world 1's floor corner, one
    particle of world 1 thrown out of the escape box; 60 steps, every step compared with the oracle."""
    import torch

    from alloy_amd.sph import scenes
    from alloy_amd.sph.config import DFSPHConfig

    cfg = DFSPHConfig(particle_radius=1.0e-3, world_bounds=((-0.004, -0.004, -0.004), (0.012, 0.012, 0.060)))
    sc = scenes.water_column(4, 4, 4, worlds=2, fills=[4, 3], spares=160, headroom=20, cfg=cfg, device="cuda")
    s = sc.make_solver()
    sp = cfg.spacing
    assert s.add_emitter(0, (2 * sp, 2 * sp, 9 * sp, 0.0, 0.0, 0.0, 1.0), (2 * sp, 2 * sp, 0.4 * sp),
                         16 * sp**3 / 0.008, (0.0, 0.0, -0.25)) == 0
    assert s.add_sink(1, (0.5 * sp, 0.5 * sp, 0.5 * sp, 0.0, 0.0, 0.0, 1.0), (sp, sp, sp)) == 0
    st = sc.model.state()
    w = sc.model.particle_world.cpu().numpy()
    phase = sc.model.sph.phase.cpu().numpy()
    thrown = int(np.nonzero((w == 1) & (phase == 1))[0][-1])
    st.particle_qd[thrown] = torch.tensor([0.0, 0.0, 40.0], device="cuda")
    o = _oracle_for(s, sc)
    o.set_state(st.particle_q.cpu().numpy(), st.particle_qd.cpu().numpy())
    _log(f"[emitters] pour/sink/escape: {sc.model.particle_count} particles, 2 worlds")
    for k in range(60):
        s.step(st, st, None, None, DT)
        o.step(DT)
        if k % 10 == 9:
            _compare(s, o, st, f"step {k + 1}")
    r = s.report()
    _log(f"[emitters] mass_in {r.mass_in}, mass_out {r.mass_out}, escaped {r.escaped}")
    assert r.mass_in[0] > 0.0 and r.mass_in[1] == 0.0
    assert r.mass_out[1] > 0.0 and r.escaped[1] >= 1 and r.escaped[0] == 0
    m = float(sc.model.particle_mass[0].item())
    assert np.allclose(r.mass_in[0] / m, round(r.mass_in[0] / m), atol=1e-6)


def test_a_closed_emitter_emits_nothing_until_opened():
    from alloy_amd.sph import scenes
    from alloy_amd.sph.config import DFSPHConfig

    cfg = DFSPHConfig(particle_radius=1.0e-3, world_bounds=((-0.004, -0.004, -0.004), (0.012, 0.012, 0.060)))
    sc = scenes.water_column(4, 4, 4, worlds=1, fills=[3], spares=160, headroom=20, cfg=cfg, device="cuda")
    s = sc.make_solver()
    sp = cfg.spacing
    e = s.add_emitter(0, (2 * sp, 2 * sp, 9 * sp, 0.0, 0.0, 0.0, 1.0), (2 * sp, 2 * sp, 0.4 * sp),
                      16 * sp**3 / 0.008, (0.0, 0.0, -0.25))
    s.set_emitter_enabled(e, False)
    st = sc.model.state()
    for _ in range(30):
        s.step(st, st, None, None, DT)
    assert s.report().mass_in[0] == 0.0
    s.set_emitter_enabled(e, True)
    for _ in range(30):
        s.step(st, st, None, None, DT)
    assert s.report().mass_in[0] > 0.0 and _exact(s.mass_ledger())
    with pytest.raises(ValueError, match="does not exist"):
        s.set_emitter_enabled(5, True)


def test_factory_escape_box_retires_ejecta_and_leaves_a_resting_scene_bitwise():
    """test_escape_box.py:63-91: the factory's automatic box retires a shot particle and counts it; a scene that
    stays inside steps bit for bit as without a box."""
    import torch

    from alloy_amd.sph import escape_box, plugin
    from alloy_amd.sph.build import add_fluid_block, add_fluid_points, wall_box
    from alloy_amd.sph.config import DFSPHConfig
    from alloy_amd.sph.model import ModelBuilder, ParticleFlags

    def scene(shot):
        cfg = DFSPHConfig(particle_radius=2.0e-3)
        b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
        b.begin_world()
        wall_box(b, (0.0, 0.0, 0.095), (0.05, 0.05, 0.005))
        add_fluid_block(b, cfg, (-0.008, -0.008, 0.1), (4, 4, 4))
        if shot:
            add_fluid_points(b, cfg, np.array([[0.03, 0.0, 0.11]]), velocity=(0.0, 0.0, 8.0))
        b.end_world()
        return b.finalize("cuda")

    m = scene(True)
    sol = plugin._factory(m, {"particle_radius": 2.0e-3})
    top = sol.config.world_bounds[1][2]
    assert 0.3 < top < 1.0
    st = m.state()
    for _ in range(120):
        sol.step(st, st, None, None, DT)
    shot = m.particle_count - 1
    assert (int(m.particle_flags[shot].item()) & int(ParticleFlags.ACTIVE)) == 0
    rep = sol.report()
    assert rep.escaped.tolist() == [1]
    assert np.isclose(rep.mass_out[0], float(m.particle_mass[shot].item()), rtol=1e-6)
    assert bool((st.particle_q[:shot, 2] < 0.2).all())
    qs = []
    for kw in ({"particle_radius": 2.0e-3}, {"particle_radius": 2.0e-3, "world_bounds": None}):
        m = scene(False)
        s = plugin._factory(m, kw)
        st = m.state()
        for _ in range(40):
            s.step(st, st, None, None, DT)
        torch.cuda.synchronize()
        qs.append(st.particle_q.clone())
    assert torch.equal(qs[0], qs[1])
    assert escape_box.auto_world_bounds(scene(False)) is not None


def test_flux_energy_export_and_retire_ledger():
    """test_flux_energy.py:37-61: the exported W_ext of one emitted layer is its kinetic energy, then the ledger is
    zeroed; SolverDFSPH.retire moves mass to the ledger exactly (int64)."""
    import torch

    from alloy_amd.sph import scenes
    from alloy_amd.sph.build import add_spares
    from alloy_amd.sph.config import DFSPHConfig
    from alloy_amd.sph.model import ModelBuilder

    cfg = DFSPHConfig(particle_radius=2.0e-3, world_bounds=((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)))
    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    b.begin_world()
    add_spares(b, cfg, 64, park=(0.0, 0.0, -0.5))
    b.end_world()
    m = b.finalize("cuda")
    s = scenes.Scene(m, cfg).make_solver()
    s.add_emitter(0, (0.0, 0.0, 0.2, 0.0, 0.0, 0.0, 1.0), (0.006, 0.006, 0.002), 1.0, (0.0, 0.0, -0.5))

    class Health:
        def __init__(self):
            self.z = np.zeros(1)
            self.w = 0.0

        def state_arrays(self):
            return {"health.energy_z_ref": self.z}

        def add_external_work(self, w_ext):
            self.w += float(np.asarray(w_ext)[0])

    st = m.state()
    h = Health()
    s.step(st, st, None, None, DT)
    s.export_flux_energy(h)
    mass = float(m.particle_mass[0].item())
    n = int(s.report().mass_in[0] / mass + 0.5)
    assert n == 9 and np.isclose(h.w, n * mass * 0.5 * 0.25, rtol=1e-6)
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    s.alloy_energy(st, torch.zeros(1, dtype=torch.float64, device="cuda"), e)
    assert np.isclose(h.w, float(e.item()), rtol=1e-6)
    s.export_flux_energy(h)  # zeroed: nothing more
    assert np.isclose(h.w, n * mass * 0.5 * 0.25, rtol=1e-6)
    mask = torch.zeros(m.particle_count, dtype=torch.int32, device="cuda")
    mask[:4] = 1
    s.retire(st, mask)
    led = s.mass_ledger()
    assert _exact(led) and led["out"][0] == 4 * int(round(mass * 2**40))


def test_pour_worlds_with_swaying_bowls_match_the_oracle_bitwise():
    """Six pour worlds of the batched pour pack (spout, swaying particle-wall bowl, spares), 150 steps: a layer every
    5 steps falls into the moving bowl; every 50 steps the GPU state equals the oracle's bit for bit."""
    from alloy_amd.sph import scenes

    sc = scenes.pour_worlds(2048, seed=4, device="cuda", subset=range(0, 2048, 341))
    s = sc.make_solver()
    st = sc.model.state()
    o = _oracle_for(s, sc)
    _log(f"[emitters] pour worlds: {sc.model.world_count} worlds, {sc.model.particle_count} particles")
    for k in range(150):
        sc.drive(st, k * DT)
        o.qd[...] = st.particle_qd.cpu().numpy()  # the same wall velocities for the oracle
        s.step(st, st, None, None, DT)
        o.step(DT)
        if k % 50 == 49:
            _compare(s, o, st, f"step {k + 1}")
    r = s.report()
    _log(f"[emitters] emitted {np.round(r.mass_in / float(sc.model.particle_mass[0].item())).astype(int).tolist()}")
    assert np.all(r.mass_in > 0) and _exact(s.mass_ledger())
    assert np.any(s.pw.world_work() != 0.0)  # a swaying bowl works on the liquid


def test_dense_particle_walls_overflow_the_wall_rows_and_match_the_oracle_bitwise():
    """A bowl lined at half the liquid spacing (four layers of radius r/2, the same thickness and volume): liquid at
    the shell has far more than the 32 wall particles within h that the slot pass records per particle, so the
    reaction split walks the cells again for those (particle_walls.py:138-177). 60 steps: the GPU state and the
    walls' fixed-point impulses equal the oracle's bit for bit."""
    from alloy_amd.sph import _abi, scenes
    from alloy_amd.sph.build import PHASE_NONE, add_box_tank, add_fluid_points, bowl_points, lattice_points
    from alloy_amd.sph.config import particle_mass
    from alloy_amd.sph.model import ModelBuilder, ParticleFlags

    cfg = scenes.bench_config()
    sp, side, R = cfg.spacing, 0.2, 0.05
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    c = np.array([0.5 * side, 0.5 * side, 2.0 * sp + R])
    b.begin_world()
    add_box_tank(b, (0.0, 0.0, 0.0), (side, side, side), wall=cfg.h)
    shell = bowl_points(tuple(c), R, 0.5 * sp, layers=4)
    b.add_particles(shell, (0.0, 0.0, 0.0), particle_mass(cfg, 1000.0), 0.5 * cfg.particle_radius,
                    int(ParticleFlags.ACTIVE | ParticleFlags.PROXY), 1000.0, 1.0e-3, PHASE_NONE)
    m = int(np.ceil(R / sp)) + 1
    pts = lattice_points(tuple(c - m * sp), (2 * m, 2 * m, m), sp)
    pts = pts[(np.linalg.norm(pts - c, axis=1) < R - 0.5 * sp) & (pts[:, 2] < c[2])]
    pts = pts + np.random.default_rng(3).uniform(-1e-4, 1e-4, pts.shape)
    add_fluid_points(b, cfg, pts)
    b.end_world()
    sc = scenes.Scene(b.finalize("cuda"), cfg, [(0.0, 0.0, 0.0)], [(side, side, side)], [len(pts)])
    d = np.linalg.norm(pts[:, None, :] - shell[None, :, :], axis=2)
    assert int((d < cfg.h).sum(axis=1).max()) > 2 * 32  # rows overflow at the start, not only by chance
    s = sc.make_solver()
    st = sc.model.state()
    o = _oracle_for(s, sc)
    _log(f"[emitters] dense bowl: {len(shell)} wall particles at half spacing, {len(pts)} liquid")
    n, S = sc.model.particle_count, sc.model.shape_count
    walked = 0  # steps in which a particle whose row overflowed pushed on the bowl: the cell-walk branch ran for it
    for k in range(60):
        s.step(st, st, None, None, DT)
        o.step(DT)
        if k % 30 == 29:
            _compare(s, o, st, f"step {k + 1}")
        if k % 10 == 9:
            rows = s.buffer(_abi.BUF_PW_ROW_COUNT, np.int32, (n,))
            shape = s.buffer(_abi.BUF_SLOT_SHAPE, np.int32, (n, 4))
            imp = s.buffer(_abi.BUF_SLOT_IMPULSE, np.float32, (n, 4, 3))
            pushing = ((shape == S) & np.any(imp != 0.0, axis=2)).any(axis=1)
            walked += int(np.any(pushing & (rows == -1)))
    _log(f"[emitters] dense bowl: cell-walk reactions in {walked} of 6 sampled steps")
    assert walked > 0
    assert np.any(s.pw.particle_impulse() != 0.0)
