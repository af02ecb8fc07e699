"""Checks that pin oracle functions the other oracle tests leave open, each from what fixes it independently of how
the oracle writes it:

- XSPH (integrate.py:150-178): Monaghan's symmetric weight m_j / rho_bar_ij makes the pair terms cancel, so the
  correction conserves momentum to rounding on any velocity field, masses unequal (integrate.py:169-171); on the rest
  lattice a shear mode v_x = A sin(k z) is an eigenvector of the correction, with the closed-form factor
  1 - c sum_j V W_ij (1 - cos(k z_ij)) per step (sign and magnitude).
- The warm-start MLP's backward pass (oracle/mlp_oracle.py): central finite differences of the float64 network with
  the bf16 roundings off, every parameter tensor.
- The resting-jitter tolerance (solve_driver.py:1-21, 63-119): the reference's three cases of
  tests/solvers/sph/test_jitter_rule.py:41-77 on primed columns.
- Slot eviction in a full row (boundary_slots.py:50-102): the reference's four cases of
  tests/solvers/sph/test_boundary_slots.py:46-69.
- The A14 dam break (twins.py:236-276, tests/solvers/sph/test_laws.py:77-85): the front stays below Ritter's dry-bed
  bound Z = 1 + 2 T (within the law's 0.025), starts at the dam and surges; also on the 8,000-particle corner dam
  break of the CPU configuration.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np
import pytest

from alloy_amd.sph import _abi, scenes
from alloy_amd.sph.build import add_box_tank, add_fluid_block, add_fluid_points
from alloy_amd.sph.config import DFSPHConfig, particle_mass
from alloy_amd.sph.kernels import kernel_w_np
from alloy_amd.sph.model import ModelBuilder
from alloy_amd.sph.solver import hydrostatic_guess
from oracle import mlp_oracle
from oracle import oracle as orc
from oracle.oracle import OracleDFSPH


# ------------------------------------------------------------------------------------------------------------ XSPH
def _free_block(cfg, counts, gravity=0.0, rest_density=1000.0):
    b = ModelBuilder(gravity=(0.0, 0.0, gravity))
    b.begin_world()
    add_fluid_block(b, cfg, (0.0, 0.0, 0.0), counts, rest_density=rest_density)
    b.end_world()
    return b.finalize("cpu")


def _xsph_step(o: OracleDFSPH, dt: float) -> np.ndarray:
    """The non-pressure update's velocity change at the oracle's state (lists, slots and densities fresh)."""
    v0 = o.qd.copy()
    o.build_neighbors()
    o.boundary_slots()
    o.density_alpha()
    o.nonpressure(dt)
    return o.qd - v0


def test_xsph_conserves_momentum_with_unequal_masses():
    """Two liquids of different rest density (so different masses and rho_i != rho_j across the interface), a jittered
    lattice and a random non-uniform velocity field, no gravity and no walls: sum_i m_i dv_i vanishes to float64
    rounding, while the correction itself is far from zero."""
    cfg = DFSPHConfig(particle_radius=1.0e-3, xsph_c=0.3)
    sp = cfg.spacing
    rng = np.random.default_rng(7)
    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    b.begin_world()
    pts = np.stack(np.meshgrid(*(np.arange(8) * sp + 0.5 * sp,) * 3, indexing="ij"), -1).reshape(-1, 3)
    pts = pts + rng.uniform(-0.2 * sp, 0.2 * sp, pts.shape)
    low = pts[:, 2] < 4 * sp
    add_fluid_points(b, cfg, pts[low], rest_density=1000.0)
    add_fluid_points(b, cfg, pts[~low], rest_density=1400.0)
    b.end_world()
    m = b.finalize("cpu")
    o = OracleDFSPH(m, cfg, 64)
    o.qd[...] = rng.normal(0.0, 0.1, o.qd.shape)
    dv = _xsph_step(o, 1.0e-3)
    mass = m.particle_mass.numpy().astype(np.float64)
    assert np.unique(mass).size == 2
    p = (mass[:, None] * dv).sum(0)
    scale = (mass[:, None] * np.abs(dv)).sum(0)
    assert np.all(scale > 0.0) and np.all(np.abs(p) <= 1.0e-12 * scale), (p, scale)


def test_xsph_damps_a_shear_mode_by_its_closed_form_factor():
    """Rest lattice (ratio 1 inside, so V_ij = m / rho0 = 1 / S for every interior pair), v_x = A sin(k z), no
    gravity: inside the block dv_x = -c A sin(k z) sum_j W_ij (1 - cos(k z_ij)) / S and dv_y = dv_z = 0; the sin(k
    z_ij) part cancels between the planes +z_ij and -z_ij. The model stores float32 positions, so the lattice offsets
    and with them the factor carry a relative error of a few 1e-6."""
    cfg = DFSPHConfig(particle_radius=1.0e-3, xsph_c=0.2)
    sp, h = cfg.spacing, cfg.h
    n = (14, 14, 24)
    m = _free_block(cfg, n)
    o = OracleDFSPH(m, cfg, 64)
    k = 2.0 * math.pi / (12 * sp)
    amp = 0.05
    o.qd[:, 0] = amp * np.sin(k * o.q[:, 2])
    dv = _xsph_step(o, 1.0e-3)
    hi = np.array(n) * sp
    inner = np.all((o.q > 2.0 * h) & (o.q < hi - 2.0 * h), axis=1)
    assert inner.sum() > 200
    # closed form over the lattice offsets within h
    ax = np.arange(-3, 4) * sp
    off = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    r = np.linalg.norm(off, axis=1)
    w = kernel_w_np(r, h, "cubic")
    S = float(w.sum())
    pair = (r > 0.0) & (r < h)
    factor = cfg.xsph_c * float(np.sum(w[pair] * (1.0 - np.cos(k * off[pair, 2])))) / S
    v0 = amp * np.sin(k * o.q[inner, 2])
    assert factor > 1.0e-3
    np.testing.assert_allclose(dv[inner, 0], -factor * v0, rtol=3.0e-5, atol=1.0e-9 * amp)
    assert np.all(dv[inner, 1:] == 0.0)


def test_xsph_off_leaves_gravity_only():
    cfg = DFSPHConfig(particle_radius=1.0e-3, xsph_c=0.0)
    m = _free_block(cfg, (4, 4, 4), gravity=-9.81)
    o = OracleDFSPH(m, cfg, 64)
    o.qd[:, 0] = np.linspace(-1.0, 1.0, len(o.qd))
    dv = _xsph_step(o, 2.0e-3)
    g = float(np.float32(-9.81))  # the model's gravity is float32
    np.testing.assert_allclose(dv, np.tile([0.0, 0.0, g * 2.0e-3], (len(dv), 1)), rtol=0, atol=1e-15)


# ------------------------------------------------------------------------------------------------------------ MLP
def _loss(p, X, T, n_total):
    y, _ = mlp_oracle.forward(p, X, rounding=False)
    return float(np.sum((y[:, :2] - T) ** 2)) / (2.0 * n_total)


def test_mlp_oracle_gradient_matches_finite_differences():
    """Every tensor (W1, b1, ..., W4, b4): 24 entries each (the real inputs and outputs, not the zero padding),
    central differences at a relative step of 1e-6; analytic and numeric agree to 1e-7 of the tensor's largest
    gradient entry, and the squared-error sum matches the loss."""
    rng = np.random.default_rng(3)
    p = rng.normal(0.0, 0.3, mlp_oracle.N_PARAMS)
    X = rng.normal(0.0, 1.0, (48, 16))
    T = rng.normal(0.0, 1.0, (48, 2))
    n_total = 96
    g, se = mlp_oracle.grad(p, X, T, n_total, rounding=False)
    assert se / (2.0 * n_total) == pytest.approx(_loss(p, X, T, n_total), rel=1e-12)
    H, IN = mlp_oracle.H, mlp_oracle.IN
    tensors = {
        "W1": (mlp_oracle.OFF_W1, [(a, b) for a in range(H) for b in range(mlp_oracle.IN_USED)], IN),
        "b1": (mlp_oracle.OFF_B1, [(a, 0) for a in range(H)], 1),
        "W2": (mlp_oracle.OFF_W2, [(a, b) for a in range(H) for b in range(H)], H),
        "b2": (mlp_oracle.OFF_B2, [(a, 0) for a in range(H)], 1),
        "W3": (mlp_oracle.OFF_W3, [(a, b) for a in range(H) for b in range(H)], H),
        "b3": (mlp_oracle.OFF_B3, [(a, 0) for a in range(H)], 1),
        "W4": (mlp_oracle.OFF_W4, [(a, b) for a in range(2) for b in range(H)], H),
        "b4": (mlp_oracle.OFF_B4, [(a, 0) for a in range(2)], 1),
    }
    for name, (off, entries, stride) in tensors.items():
        pick = [entries[i] for i in rng.choice(len(entries), size=min(24, len(entries)), replace=False)]
        idx = np.array([off + a * stride + b for a, b in pick])
        num = np.empty(len(idx))
        for t, j in enumerate(idx):
            e = 1.0e-6 * max(abs(p[j]), 1.0)
            pp, pm = p.copy(), p.copy()
            pp[j] += e
            pm[j] -= e
            num[t] = (_loss(pp, X, T, n_total) - _loss(pm, X, T, n_total)) / (2.0 * e)
        ref = np.abs(g[idx]).max()
        assert ref > 0.0, name
        assert np.abs(num - g[idx]).max() <= 1.0e-7 * ref, (name, np.abs(num - g[idx]).max() / ref)
    # the padding carries no gradient: W1 input columns 13..15, W4 rows and b4 entries 2..15
    W1 = g[mlp_oracle.OFF_W1:mlp_oracle.OFF_B1].reshape(H, IN)
    assert np.all(W1[:, mlp_oracle.IN_USED:] == 0.0)
    assert np.all(g[mlp_oracle.OFF_W4 + 2 * H:mlp_oracle.OFF_B4] == 0.0) and np.all(g[mlp_oracle.OFF_B4 + 2:] == 0.0)


def test_mlp_oracle_rounding_only_moves_the_gradient_slightly():
    """The bf16 roundings (the GPU's MFMA operands) perturb the float64 gradient at the percent level, no more."""
    rng = np.random.default_rng(4)
    p = rng.normal(0.0, 0.3, mlp_oracle.N_PARAMS)
    X, T = rng.normal(size=(64, 16)), rng.normal(size=(64, 2))
    g_r, _ = mlp_oracle.grad(p, X, T)
    g_x, _ = mlp_oracle.grad(p, X, T, rounding=False)
    assert np.linalg.norm(g_r - g_x) <= 0.05 * np.linalg.norm(g_x)


# ----------------------------------------------------------------------------------------------- resting-jitter rule
JITTER_DT = 1.25e-3


def _columns(cfg: DFSPHConfig, depths, gravity: float = -9.81, settle: int = 20):
    """test_jitter_rule.py:21-37: a 6 x 6 column per world in a box tank (walls h thick, 6 spacings of headroom),
    primed, then ``settle`` steps."""
    sp = cfg.spacing
    b = ModelBuilder(gravity=(0.0, 0.0, gravity))
    for d in depths:
        b.begin_world()
        add_box_tank(b, (0.0, 0.0, 0.0), (6 * sp, 6 * sp, (d + 6) * sp), wall=cfg.h)
        add_fluid_block(b, cfg, (0.0, 0.0, 0.0), (6, 6, d))
        b.end_world()
    m = b.finalize("cpu")
    o = OracleDFSPH(m, cfg, 32)
    o.prime(hydrostatic_guess(m, m.state(), sp), JITTER_DT)
    for _ in range(settle):
        o.step(JITTER_DT)
    return o


def _density_tol(o: OracleDFSPH, cfg: DFSPHConfig) -> np.ndarray:
    return np.asarray(o.buffer(_abi.BUF_TOL_SCALE)[: o.W], np.float64) * cfg.max_density_error


def test_jitter_tolerance_follows_the_depth_per_world():
    """test_jitter_rule.py:41-46: without the flow allowance, tol = rest_jitter_speed dt / H per world, H the
    hydrostatic depth (within 15 %: the head of a settled column is about 1.08 depths)."""
    cfg = DFSPHConfig(particle_radius=2.0e-3, rest_jitter_flow_ratio=0.0)
    o = _columns(cfg, (8, 25))
    expected = cfg.rest_jitter_speed * JITTER_DT / (np.array([8, 25]) * cfg.spacing)
    np.testing.assert_allclose(_density_tol(o, cfg), expected, rtol=0.15)
    assert np.all(expected < cfg.max_density_error)


def test_jitter_flow_allowance_uses_the_mean_velocity_before_the_step():
    """test_jitter_rule.py:49-68: a sideways kick of the whole column; v = max(rest_jitter_speed, ratio U) with U the
    liquid's mean velocity in the state the step starts from, H = 1.5 sum k^2 / (|g| sum k) from the stored density
    warm start k = kd / dt^2."""
    cfg = DFSPHConfig(particle_radius=2.0e-3)
    o = _columns(cfg, (8, 25))
    n = o.n
    liquid = o.model.sph.phase.numpy()[:n] >= 1
    world = o.model.particle_world.numpy()[:n]
    o.qd[liquid, 0] += 0.02
    v = o.qd.astype(np.float64)
    u = np.array([np.linalg.norm(v[liquid & (world == w)].mean(0)) for w in (0, 1)])
    kd = np.maximum(o.buffer(_abi.BUF_KD)[:n].astype(np.float64), 0.0)
    depth = np.array([1.5 * np.sum(kd[liquid & (world == w)] ** 2) /
                      (np.sum(kd[liquid & (world == w)]) * JITTER_DT**2 * 9.81) for w in (0, 1)])
    o.step(JITTER_DT)
    expected = np.minimum(cfg.max_density_error,
                          np.maximum(cfg.rest_jitter_speed, cfg.rest_jitter_flow_ratio * u) * JITTER_DT / depth)
    assert np.all(u > 0.015) and np.all(expected < cfg.max_density_error), (u, expected)
    assert np.all(cfg.rest_jitter_flow_ratio * u > cfg.rest_jitter_speed)  # the allowance, not the floor speed
    np.testing.assert_allclose(_density_tol(o, cfg), expected, rtol=1e-3)


def test_jitter_rule_off_without_gravity_with_none_and_floored():
    """test_jitter_rule.py:71-77."""
    cfg = DFSPHConfig(particle_radius=2.0e-3)
    assert _density_tol(_columns(cfg, (8,), gravity=0.0, settle=3), cfg)[0] == pytest.approx(cfg.max_density_error)
    off = dataclasses.replace(cfg, rest_jitter_speed=None)
    assert _density_tol(_columns(off, (8,), settle=3), off)[0] == pytest.approx(cfg.max_density_error)
    floor = dataclasses.replace(cfg, min_density_error=2.0e-4)
    assert _density_tol(_columns(floor, (8,)), floor)[0] == pytest.approx(2.0e-4)


# ----------------------------------------------------------------------------------------------- slot eviction
NEAR = 0.4e-3  # 0.1 h at 2 mm particles
FLOOR = 99
UP = (0.0, 0.0, 1.0)


def _wall(deg: float):
    a = math.radians(deg)
    return (math.cos(a), math.sin(a), 0.0)


def test_slot_eviction_keeps_the_nearer_slab_at_an_inner_vertex():
    """test_boundary_slots.py:46-49: the farthest duplicate (60 deg, facing the 45 deg slab's way) goes."""
    slots = [(FLOOR, 0.0, UP), (2, 1.0e-6, _wall(30)), (1, 1.67e-3, _wall(15)), (4, 1.88e-3, _wall(60))]
    kept, after, phi = orc.shape_insert(slots, (3, 1.0e-4, _wall(45)), NEAR)
    assert kept == 1 and after == [FLOOR, 2, 3, 1] and np.all(np.diff(phi) >= 0.0)


def test_slot_eviction_lets_the_floor_take_a_corner_slot_from_four_slabs():
    """test_boundary_slots.py:52-55."""
    slots = [(2, 1.0e-4, _wall(30)), (3, 2.0e-4, _wall(45)), (1, 5.0e-4, _wall(15)), (4, 6.0e-4, _wall(60))]
    kept, after, phi = orc.shape_insert(slots, (FLOOR, 1.0e-3, UP), NEAR)
    assert kept == 1 and after == [2, 3, 1, FLOOR] and np.all(np.diff(phi) >= 0.0)


def test_slot_eviction_drops_a_farther_duplicate():
    """test_boundary_slots.py:58-61."""
    slots = [(FLOOR, 0.0, UP), (2, 1.0e-6, _wall(30)), (3, 1.0e-4, _wall(45)), (1, 1.67e-3, _wall(15))]
    kept, after, phi = orc.shape_insert(slots, (4, 1.88e-3, _wall(60)), NEAR)
    assert kept == 0 and after == [FLOOR, 2, 3, 1] and np.all(np.diff(phi) >= 0.0)


@pytest.mark.parametrize("deg", [90.0, 180.0])
def test_slot_eviction_never_evicts_walls_within_a_cfl_step(deg):
    """test_boundary_slots.py:64-69: four slots within ``near`` stay; a farther wall facing another way finds no
    room."""
    slots = [(FLOOR, 0.0, UP), (2, 0.0, _wall(30)), (3, 1.0e-4, _wall(45)), (1, 2.0e-4, _wall(15))]
    kept, after, phi = orc.shape_insert(slots, (7, 3.0e-3, _wall(deg)), NEAR)
    assert kept == 0 and after == [FLOOR, 2, 3, 1] and np.all(np.diff(phi) >= 0.0)


@pytest.mark.parametrize("apart,after", [(25.0, [FLOOR, 2, 3, 1]), (35.0, [FLOOR, 2, 4, 3])])
def test_slot_eviction_same_facing_threshold(apart, after):
    """COS_SAME = 0.85 (boundary_slots.py:9, "within about 32 deg"): a slab 25 deg off a nearer one is its duplicate
    and makes room although a farther wall remains (cos 25 deg = 0.906); at 35 deg apart (cos = 0.819) it is a wall of
    its own and the farthest slot goes instead."""
    slots = [(FLOOR, 0.0, UP), (2, 1.0e-6, _wall(0.0)), (4, 1.0e-3, _wall(apart)), (1, 1.88e-3, _wall(90.0))]
    kept, got, phi = orc.shape_insert(slots, (3, 1.2e-3, _wall(180.0)), NEAR)
    assert kept == 1 and got == after and np.all(np.diff(phi) >= 0.0)


def test_slot_insert_below_four_keeps_phi_order():
    """A row with room inserts in phi order and never evicts."""
    slots = [(FLOOR, 1.0e-4, UP), (2, 9.0e-4, _wall(30))]
    kept, after, phi = orc.shape_insert(slots, (5, 5.0e-4, _wall(30)), NEAR)
    assert kept == 1 and after == [FLOOR, 5, 2, -1] and phi[:3].tolist() == pytest.approx([1e-4, 5e-4, 9e-4])


# ----------------------------------------------------------------------------------------------- dam-break front
def _front(o: OracleDFSPH, r: float) -> float:
    liquid = o.model.sph.phase.numpy()[: o.n] >= 1
    x = o.q[liquid]
    return float(x[x[:, 2] < 4.0 * r, 0].max()) + r  # probes.front_kernel + store_max_kernel's + r


def _ritter(o: OracleDFSPH, cfg: DFSPHConfig, a_m: float, dt: float, t_end: float, every: int):
    T, Z = [0.0], [_front(o, cfg.particle_radius) / a_m]
    for k in range(1, int(round(t_end / dt)) + 1):
        o.step(dt)
        if k % every == 0:
            T.append(k * dt * math.sqrt(2.0 * 9.81 / a_m))
            Z.append(_front(o, cfg.particle_radius) / a_m)
    return np.array(T), np.array(Z)


def test_a14_dam_break_front_stays_below_ritter():
    """A14 quick (twins.py:236-276, test_laws.py:77-85): an a x depth x 2a column (a = 12, depth 3 spacings at
    r = 1 mm) against the x_lo wall of a 5a x depth x (2a + 10) tank, dt 78 us, 0.11 s; Z = front / a,
    T = t sqrt(2 g / a): Z(0) = 1, Z never recedes, Z - (1 + 2 T) <= 0.025 for T <= 3, and Z > 2 by T = 3."""
    cfg = DFSPHConfig(particle_radius=1.0e-3)
    sp = cfg.spacing
    a, depth = 12, 3
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    b.begin_world()
    add_box_tank(b, (0.0, 0.0, 0.0), (5 * a * sp, depth * sp, (2 * a + 10) * sp), wall=cfg.h)
    add_fluid_block(b, cfg, (0.0, 0.0, 0.0), (a, depth, 2 * a))
    b.end_world()
    o = OracleDFSPH(b.finalize("cpu"), cfg, 32)
    T, Z = _ritter(o, cfg, a * sp, 0.02 / 256, 0.11, 32)
    within = T <= 3.0
    assert Z[0] == pytest.approx(1.0, abs=0.05) and np.all(np.diff(Z) >= -1e-3), Z
    assert np.max(Z[within] - (1.0 + 2.0 * T[within])) <= 0.025, np.max(Z[within] - (1.0 + 2.0 * T[within]))
    assert Z[within][-1] > 2.0


def test_cube_dam_break_front_stays_below_ritter():
    """The 8,000-particle dam break (20^3 lattice at 0.01 m, jitter 1e-4 m, in a corner of a 0.4 m tank), dt 1 ms,
    200 steps (T = 2): a corner collapse with its height equal to its width, so its front is slower than the 2:1
    column's and the same bound holds. The front advances until it meets the far wall at Z = 2, which it reaches."""
    sc = scenes.dam_break(20, 0.4, 0, device="cpu")
    o = OracleDFSPH(sc.model, sc.config, 32)
    T, Z = _ritter(o, sc.config, 20 * sc.config.spacing, scenes.DT, 0.2, 10)
    free = np.arange(len(Z)) < np.argmax(Z >= 1.95) if np.any(Z >= 1.95) else np.ones(len(Z), bool)
    assert Z[0] == pytest.approx(1.0, abs=0.05) and np.all(np.diff(Z[free]) >= -1e-3), Z
    assert np.max(Z - (1.0 + 2.0 * T)) <= 0.025
    assert Z.max() >= 1.95
