"""The oracle's emitter pass (emitters.py:67-172, runtime.py:197-347) on the CPU: the exact int64 mass ledger with an
emitter, a sink and the escape box (tests/solvers/sph/test_mass.py:13-41 of the reference), a closed emitter
(test_mass.py:44-62), the escape box retiring ejecta (test_escape_box.py:63-75), the energy flux against closed forms
(test_flux_energy.py:37-86), and the host bookkeeping (layer lattice, pool split)."""

from __future__ import annotations

import numpy as np
import pytest

from alloy_amd.sph import _abi, scenes
from alloy_amd.sph.build import add_fluid_points, add_spares
from alloy_amd.sph.config import DFSPHConfig
from alloy_amd.sph.emitters import EmitterManager, EmitterTable, layer_offsets, split_pools
from alloy_amd.sph.model import ModelBuilder, ParticleFlags
from oracle.oracle import OracleDFSPH

DT = 1.0e-3


def _oracle_with_emitters(sc):
    o = OracleDFSPH(sc.model, sc.config, 32)

    def cur():
        c = np.asarray(o.buffer(_abi.BUF_EMIT_CURSOR)).copy()
        return np.concatenate([c, np.zeros(len(mgr.table.world) - len(c), np.int32)])

    mgr = EmitterManager(sc.model, sc.config, o.set_emitters, cur)
    if mgr.active:
        mgr.upload()
    return o, mgr


def _exact(led):
    return np.array_equal(led["active"] + led["out"], led["initial"] + led["in"])


def test_layer_offsets_whole_columns():
    # a nozzle 0.3 m wide at 0.1 m is 3 columns (0.3 / 0.1 = 2.9999999999999996 without the slack), emitters.py:48-50
    offs = layer_offsets((0.15, 0.15, 0.01), (0.0, 0.0, -1.0), 0.1)
    assert offs.shape == (9, 3)
    assert np.allclose(np.unique(offs[:, 0]), [-0.1, 0.0, 0.1]) and np.all(offs[:, 2] == 0.0)
    # the cross-section is perpendicular to the velocity's dominant axis; no velocity: z
    offs = layer_offsets((0.02, 0.05, 0.05), (3.0, 0.5, 0.0), 0.01)
    assert offs.shape == (100, 3) and np.all(offs[:, 0] == 0.0)
    assert layer_offsets((0.001, 0.001, 0.001), (0.0, 0.0, 0.0), 0.01).shape == (1, 3)


def test_split_pools_shares_spares_and_keeps_an_owned_pool():
    t = EmitterTable(world=[0, 0, 0], own=[False, True, False], pool=[np.zeros(0, np.int32),
                                                                       np.array([3, 4], np.int32),
                                                                       np.zeros(0, np.int32)])
    split_pools(t, 0, np.arange(10, dtype=np.int32), np.zeros(3, np.int32))
    assert t.pool[1].tolist() == [3, 4]
    assert t.pool[0].tolist() == [0, 1, 2, 5] and t.pool[2].tolist() == [6, 7, 8, 9]
    # after emitter 0 has emitted, only the last one's share is recomputed
    split_pools(t, 0, np.arange(12, dtype=np.int32), np.array([2, 0, 0], np.int32))
    assert t.pool[0].tolist() == [0, 1, 2, 5] and t.pool[2].tolist() == [6, 7, 8, 9, 10, 11]


def test_ledger_exact_with_emitter_sink_and_escape_box():
    cfg = DFSPHConfig(particle_radius=1.0e-3, world_bounds=((-0.004, -0.004, -0.004), (0.012, 0.012, 0.060)))
    sc = scenes.water_column(4, 4, 4, worlds=2, fills=[4, 3], spares=160, headroom=20, cfg=cfg)
    sp = cfg.spacing
    o, mgr = _oracle_with_emitters(sc)
    e = mgr.add_emitter(0, (2 * sp, 2 * sp, 9 * sp, 0.0, 0.0, 0.0, 1.0), (2 * sp, 2 * sp, 0.4 * sp),
                        16 * sp**3 / 0.008, (0.0, 0.0, -0.25))
    k = mgr.add_sink(1, (0.5 * sp, 0.5 * sp, 0.5 * sp, 0.0, 0.0, 0.0, 1.0), (sp, sp, sp))
    assert (e, k) == (0, 0)
    w = sc.model.particle_world.numpy()
    phase = sc.model.sph.phase.numpy()
    thrown = int(np.nonzero((w == 1) & (phase == 1))[0][-1])
    o.qd[thrown] = (0.0, 0.0, 40.0)
    led0 = o.mass_ledger()
    assert _exact(led0) and np.all(led0["in"] == 0)
    for _ in range(50):
        o.step(DT)
    led = o.mass_ledger()
    assert _exact(led)
    assert led["in"][0] > 0 and led["in"][1] == 0
    assert led["out"][1] > 0
    esc = o.buffer(_abi.BUF_ESCAPED)
    assert esc[1] >= 1 and esc[0] == 0
    m_fixed = int(np.round(float(np.float32(sc.model.particle_mass[0].item())) * 2.0**40))
    assert led["in"][0] % m_fixed == 0  # whole particles emitted
    n_emitted = int(led["in"][0] // m_fixed)
    flags, ph = o.buffer(_abi.BUF_FLAGS), o.buffer(_abi.BUF_PHASE)
    spares0 = np.nonzero((w == 0) & (phase == 0))[0]
    now_liquid = spares0[(ph[spares0] == 1) & ((flags[spares0] & int(ParticleFlags.ACTIVE)) != 0)]
    assert len(now_liquid) == n_emitted and n_emitted % 16 == 0  # layers of 4 x 4
    assert o.buffer(_abi.BUF_EMIT_CURSOR)[0] == n_emitted
    # the thrown particle retired: phase -2, ACTIVE cleared, velocity zeroed
    assert ph[thrown] == -2 and (flags[thrown] & int(ParticleFlags.ACTIVE)) == 0 and np.all(o.qd[thrown] == 0)


def test_a_closed_emitter_emits_nothing_until_opened():
    cfg = DFSPHConfig(particle_radius=1.0e-3, world_bounds=((-0.004, -0.004, -0.004), (0.012, 0.012, 0.060)))
    sc = scenes.water_column(4, 4, 4, worlds=1, fills=[3], spares=160, headroom=20, cfg=cfg)
    sp = cfg.spacing
    o, mgr = _oracle_with_emitters(sc)
    e = mgr.add_emitter(0, (2 * sp, 2 * sp, 9 * sp, 0.0, 0.0, 0.0, 1.0), (2 * sp, 2 * sp, 0.4 * sp),
                        16 * sp**3 / 0.008, (0.0, 0.0, -0.25))
    o.set_emitter_enabled(e, False)
    for _ in range(30):
        o.step(DT)
    assert o.mass_ledger()["in"][0] == 0 and float(o.buffer(_abi.BUF_EMIT_T)[0]) == 0.0
    o.set_emitter_enabled(e, True)
    for _ in range(30):
        o.step(DT)
    led = o.mass_ledger()
    assert led["in"][0] > 0 and _exact(led)


def _lone_particles(points, velocity, gravity, bounds, spares=0):
    cfg = scenes.bench_config(particle_radius=0.002, world_bounds=bounds)
    b = ModelBuilder(gravity=gravity)
    b.begin_world()
    if len(points):
        add_fluid_points(b, cfg, np.asarray(points), velocity=velocity)
    if spares:
        add_spares(b, cfg, spares, park=(0.0, 0.0, -0.5))
    b.end_world()
    return scenes.Scene(b.finalize("cpu"), cfg)


def test_escape_box_retires_ejecta_and_counts_them():
    sc = _lone_particles([[0.0, 0.0, 0.0], [0.05, 0.0, 0.0]], (0.0, 0.0, 0.0), (0.0, 0.0, -9.81),
                         ((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1)))
    o, _ = _oracle_with_emitters(sc)
    o.qd[1] = (0.0, 0.0, 8.0)  # 8 m/s up: leaves the box after ~13 steps
    for _ in range(30):
        o.step(DT)
    assert o.buffer(_abi.BUF_ESCAPED).tolist() == [1]
    led = o.mass_ledger()
    assert _exact(led) and led["out"][0] == int(np.round(float(np.float32(sc.model.particle_mass[1].item())) * 2**40))
    assert (o.buffer(_abi.BUF_FLAGS)[0] & int(ParticleFlags.ACTIVE)) != 0


def test_emitted_energy_is_external_work():
    """Zero g: one layer emitted at 0.5 m/s brings in n m v^2 / 2 (test_flux_energy.py:37-61)."""
    sc = _lone_particles([], (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)), spares=64)
    o, mgr = _oracle_with_emitters(sc)
    mgr.add_emitter(0, (0.0, 0.0, 0.2, 0.0, 0.0, 0.0, 1.0), (0.006, 0.006, 0.002), 1.0, (0.0, 0.0, -0.5))
    o.step(DT)
    m = float(np.float32(sc.model.particle_mass[0].item()))
    n = int(o.mass_ledger()["in"][0] // int(np.round(m * 2**40)))
    assert n == 9  # 3 x 3 layer at 4 mm
    w_ext = float(o.buffer(_abi.BUF_FLUX_E)[0]) / 2**50
    assert np.isclose(w_ext, n * m * 0.5 * 0.25, rtol=1e-6)
    assert o.buffer(_abi.BUF_FLUX_G)[0] == 0


def test_retired_energy_leaves_at_the_monitors_z_ref():
    """Gravity on: a lone particle outside the box retires after its step; W_ext = -m (v^2/2 + |g| (z - z_ref)) of
    its end-of-step state (one symplectic Euler step of free fall), for any z_ref."""
    v0, z0, g = np.array([0.3, 0.0, 1.2]), 0.15, 9.81
    sc = _lone_particles([[0.0, 0.0, 0.0], [0.0, 0.0, z0]], (0.3, 0.0, 1.2), (0.0, 0.0, -g),
                         ((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1)))
    o, _ = _oracle_with_emitters(sc)
    o.step(DT)
    assert o.buffer(_abi.BUF_ESCAPED).tolist() == [1]
    m = float(np.float32(sc.model.particle_mass[1].item()))
    v1 = v0 + np.array([0.0, 0.0, -g]) * DT
    z1 = z0 + v1[2] * DT
    fe, fg = float(o.buffer(_abi.BUF_FLUX_E)[0]) / 2**50, float(o.buffer(_abi.BUF_FLUX_G)[0]) / 2**40
    for z_ref in (0.0, 0.05):
        w_ext = fe + fg * z_ref
        assert np.isclose(w_ext, -m * (0.5 * v1 @ v1 + g * (z1 - z_ref)), rtol=1e-5)


def test_layer_emission_positions_and_pool_order():
    """A layer lands on the lattice through the region transform (a quarter turn about z), spares in pool order."""
    sc = _lone_particles([], (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), None, spares=20)
    o, mgr = _oracle_with_emitters(sc)
    q = (0.0, 0.0, np.sin(np.pi / 4), np.cos(np.pi / 4))  # 90 degrees about z
    mgr.add_emitter(0, (0.1, 0.2, 0.3, *q), (0.004, 0.002, 0.001), 1.0, (0.0, 0.0, -1.0))
    o.step(DT)
    flags = o.buffer(_abi.BUF_FLAGS)
    active = np.nonzero((flags & int(ParticleFlags.ACTIVE)) != 0)[0]
    assert active.tolist() == [0, 1]  # a 2 x 1 layer from the first spares
    offs = layer_offsets((0.004, 0.002, 0.001), (0.0, 0.0, -1.0), 0.004)
    want = np.array([[0.1 - o_[1], 0.2 + o_[0], 0.3] for o_ in offs])  # (x, y) -> (-y, x)
    assert np.allclose(o.q[active], want, atol=1e-6)  # emitted at the end of the step: not yet advected
    assert np.allclose(o.qd[active], [[0.0, 0.0, -1.0]] * 2)


def test_bad_emitter_arguments_raise():
    sc = _lone_particles([[0.0, 0.0, 0.0]], (0.0, 0.0, 0.0), (0.0, 0.0, -9.81), None, spares=4)
    o, mgr = _oracle_with_emitters(sc)
    with pytest.raises(ValueError, match="out of range"):
        mgr.add_emitter(3, (0, 0, 0, 0, 0, 0, 1), (0.01, 0.01, 0.01), 1.0, (0, 0, -1))
    with pytest.raises(ValueError, match="rate"):
        mgr.add_emitter(0, (0, 0, 0, 0, 0, 0, 1), (0.01, 0.01, 0.01), 0.0, (0, 0, -1))
    with pytest.raises(ValueError, match="out of range"):
        mgr.add_sink(-1, (0, 0, 0, 0, 0, 0, 1), (0.01, 0.01, 0.01))


def _fixed_impulse_of_wall_slots(o, S):
    shape = o.buffer(_abi.BUF_SLOT_SHAPE)[: 4 * o.n].reshape(o.n, 4)
    imp = o.buffer(_abi.BUF_SLOT_IMPULSE)[: 12 * o.n].reshape(o.n, 4, 3).astype(np.float64)
    sel = shape == S
    x = imp[sel] * 2.0**50  # exact in float64; rounded half away from zero, as llround and Warp's round do
    return (np.sign(x) * np.floor(np.abs(x) + 0.5)).astype(np.int64).sum(axis=0), int(sel.sum())


def test_particle_wall_reactions_conserve_each_slot_impulse_exactly():
    """particle_walls.py:138-177: every step, the parts handed to the wall particles (plus the rounding remainder
    given to the largest share) sum to exactly the fixed-point impulse of the particle-wall slots; the window sums
    over steps; one-way coupling hands nothing to the walls."""
    import dataclasses

    sc = scenes.bowl_pool(1)
    o = OracleDFSPH(sc.model, sc.config, 32)
    S = sc.model.shape_count
    total = np.zeros(3, np.int64)
    slots = 0
    for _ in range(8):
        o.step(DT)
        want, k = _fixed_impulse_of_wall_slots(o, S)
        total += want
        slots += k
    exch = o.buffer(_abi.BUF_PW_EXCH).reshape(-1, 3)
    assert slots > 0 and np.any(total != 0)
    assert np.array_equal(exch.sum(axis=0), total)
    n_wall = int(np.count_nonzero(o.tables.wall_volume > 0))
    assert exch.shape[0] == n_wall and np.count_nonzero(exch.any(axis=1)) > 1
    assert np.all(o.buffer(_abi.BUF_PW_WORK) == 0)  # a static bowl does no work
    sc1 = scenes.bowl_pool(1, cfg=dataclasses.replace(sc.config, coupling="one_way"))
    o1 = OracleDFSPH(sc1.model, sc1.config, 32)
    for _ in range(4):
        o1.step(DT)
    assert not o1.buffer(_abi.BUF_PW_EXCH).any()
