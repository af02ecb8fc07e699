"""The oracle's rarely taken branches, each against the reference's own law for it:

- compression relief (pressure.py:44-56, 257-258; tests/solvers/sph/test_compression_relief.py): a layer pressed into
  the floor's projection skin leaves below the relief speed, and is shot out uncapped; in a resting column the cap is
  never reached, so relief on and off give the same bits;
- the lone-particle contact limit (pressure.py:393-414): a lone particle pushed off a wall by the density solve leaves
  no faster than contact_separation_speed, and the limit is what binds;
- lone-particle wall friction (pressure.py:417-459; tests/solvers/sph/test_wall_friction_xsph.py:34-51): a particle
  sliding on a floor slows at mu g and stops, and keeps its speed without friction;
- the fast-particle path sweep (integrate.py:38-114; tests/solvers/sph/test_thin_walls.py:78-101): a step past a
  1 mm wall that the boundary slots do not see ends before the wall.
"""

from __future__ import annotations

import numpy as np
import pytest

from alloy_amd.sph import _abi
from alloy_amd.sph.build import add_box_tank, add_fluid_points, wall_box
from alloy_amd.sph.config import DFSPHConfig
from alloy_amd.sph.model import ModelBuilder
from alloy_amd.sph.solver import hydrostatic_guess
from oracle.oracle import OracleDFSPH


# ------------------------------------------------------------------------------------------------ compression relief
RELIEF_DT = 1.3e-4  # the step a CFL controller picks right after an impact at a few m/s (r = 1 mm)


def _pressed_column(relief: float | None, steps: int):
    """test_compression_relief.py:21-48: a 6 x 6 x 8 column (r = 1 mm) two spacings clear of the side walls, its bottom
    layer pressed into the floor's projection skin (z = 1 um); (first step's max density ratio, max speed over the
    steps, final positions)."""
    cfg = DFSPHConfig(particle_radius=1.0e-3, compression_relief_speed=relief)
    sp = cfg.spacing
    n, gap = 6, 2
    w = (n + 2 * gap) * sp
    g = (np.arange(n) + 0.5 + gap) * sp
    pts = np.array([(x, y, (k + 0.5) * sp) for k in range(8) for y in g for x in g])
    pts[pts[:, 2] < sp, 2] = 1.0e-6
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    b.begin_world()
    add_box_tank(b, (0.0, 0.0, 0.0), (w, w, 18 * sp), wall=cfg.h)
    add_fluid_points(b, cfg, pts)
    b.end_world()
    o = OracleDFSPH(b.finalize("cpu"), cfg, 32)
    ratio0, vmax = 0.0, 0.0
    for k in range(steps):
        o.step(RELIEF_DT)
        if k == 0:
            ratio0 = float(o.buffer(_abi.BUF_RATIO)[: o.n].max())
        vmax = max(vmax, float(np.linalg.norm(o.qd, axis=1).max()))
    return ratio0, vmax, o.q.copy()


def test_skin_particles_leave_at_the_relief_speed():
    """test_compression_relief.py:51-59: capped, the pressed layer (ratio above 2) lifts out below the relief speed
    and stays above the floor; uncapped, it is shot out at several m/s within two steps."""
    ratio0, vmax, q = _pressed_column(0.5, 30)
    assert ratio0 > 2.0, ratio0
    assert vmax < 0.5, vmax
    assert np.all(np.isfinite(q)) and q[:, 2].min() > 0.0
    _, vmax_off, _ = _pressed_column(None, 2)
    assert vmax_off > 2.0, vmax_off


def test_relief_is_inactive_in_a_resting_column():
    """test_compression_relief.py:62-83: a primed resting column never reaches the cap, so relief 0.5 and relief off
    give the same bits over 20 steps."""
    out = []
    for relief in (0.5, None):
        cfg = DFSPHConfig(particle_radius=2.0e-3, compression_relief_speed=relief)
        sp = cfg.spacing
        b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
        b.begin_world()
        add_box_tank(b, (0.0, 0.0, 0.0), (6 * sp, 6 * sp, 14 * sp), wall=cfg.h)
        g = (np.arange(6) + 0.5) * sp
        add_fluid_points(b, cfg, np.array([(x, y, (k + 0.5) * sp) for k in range(8) for y in g for x in g]))
        b.end_world()
        m = b.finalize("cpu")
        o = OracleDFSPH(m, cfg, 32)
        o.prime(hydrostatic_guess(m, m.state(), sp), 1.25e-3)
        for _ in range(20):
            o.step(1.25e-3)
        out.append((o.q.copy(), o.qd.copy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


# ------------------------------------------------------------------------------------------------ lone particles
def _lone(z: float, velocity, cfg: DFSPHConfig, gravity: float = -9.81, mu: float = 1.0):
    """One liquid particle above a static 2 x 2 m floor box (top face z = 0), friction ``mu``
    (test_wall_friction_xsph.py:21-31)."""
    b = ModelBuilder(gravity=(0.0, 0.0, gravity))
    b.begin_world()
    wall_box(b, (0.0, 0.0, -0.005), (2.0, 2.0, 0.005), mu=mu)
    add_fluid_points(b, cfg, np.array([[0.0, 0.0, z]]), velocity=velocity)
    b.end_world()
    return b.finalize("cpu")


@pytest.mark.parametrize("v_sep", [0.2, 0.05])
def test_a_lone_particle_leaves_a_wall_no_faster_than_the_separation_speed(v_sep):
    """A lone particle at rest 0.1 r above the floor, no gravity: the density solve pushes it off the wall (at about
    0.45 m/s unlimited), and the contact limit holds its separation speed to exactly v_sep after the solve
    (pressure.py:393-414). Relief off, so the limit alone binds: with a separation speed of 100 m/s the same solve
    pushes it faster."""
    speeds = {}
    for vs in (v_sep, 100.0):
        cfg = DFSPHConfig(particle_radius=2.0e-3, contact_separation_speed=vs, compression_relief_speed=None)
        o = OracleDFSPH(_lone(0.2e-3, (0.0, 0.0, 0.0), cfg, gravity=0.0), cfg, 64)
        dt = 1.25e-3
        o.build_neighbors()
        o.boundary_slots()
        o.density_alpha()
        assert o.buffer(_abi.BUF_COUNT)[0] == 0 and o.buffer(_abi.BUF_RATIO)[0] > 1.0  # lone, compressed by the wall
        o.solve(1, dt)
        speeds[vs] = float(o.qd[0, 2])
    assert 0.0 < speeds[v_sep] <= v_sep * (1.0 + 1e-9), speeds
    assert speeds[v_sep] == pytest.approx(v_sep, rel=1e-9)  # the cap binds exactly: u + dt k |G| = v_sep
    assert speeds[100.0] > v_sep


@pytest.mark.parametrize("friction", [True, False])
def test_lone_particle_on_a_floor_slows_at_mu_g(friction):
    """test_wall_friction_xsph.py:34-51: a particle 2 mm above the floor sliding at 2.6 m/s (r = 2 mm, dt 1.25 ms,
    320 steps) lands, is held off the floor by the contact pressure, and slows at mu g (mu 1, a stop after 0.27 s) to
    a stop; without wall friction it keeps 2.6 m/s."""
    cfg = DFSPHConfig(particle_radius=2.0e-3, wall_friction=friction)
    dt, mu = 1.25e-3, 1.0
    o = OracleDFSPH(_lone(0.002, (2.6, 0.0, 0.0), cfg, mu=mu), cfg, 32)
    vx = []
    for _ in range(320):
        o.step(dt)
        vx.append(float(o.qd[0, 0]))
    t = dt * np.arange(1, 321)
    if friction:
        expect = np.maximum(2.6 - mu * 9.81 * t, 0.0)
        late = t >= 0.05  # it falls the 2 mm onto the floor first; the landing's normal impulse makes up mu g t0
        assert np.abs(np.array(vx)[late] - expect[late]).max() < 0.01
        assert vx[-1] == 0.0
    else:
        assert np.allclose(vx, 2.6)
    assert o.q[0, 2] > 0.0  # never through the floor


# ------------------------------------------------------------------------------------------------ path sweep
SWEEP_DT = 0.02 / 64


def _channel(step: float):
    """test_thin_walls.py:78-89: one particle (r = 2 mm, h = 8 mm, SWEEP h = 4 mm) at the origin moving +z by ``step``
    per step, in a channel of four 10 mm walls 2 mm off that fill its 4 boundary slots, so a 1 mm box wall 2.5 mm
    ahead is not in them: only the path sweep, measured against the 1 mm wall, can stop it. No gravity."""
    cfg = DFSPHConfig(particle_radius=2.0e-3)
    b = ModelBuilder(gravity=(0.0, 0.0, 0.0))
    b.begin_world()
    for x, y in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        wall_box(b, (x * 7.0e-3, y * 7.0e-3, 0.0), (5.0e-3 if x else 7.0e-3, 5.0e-3 if y else 7.0e-3, 5.0e-3))
    wall_box(b, (0.0, 0.0, 2.5e-3 + 0.5e-3), (0.02, 0.02, 0.5e-3))
    add_fluid_points(b, cfg, np.zeros((1, 3)), velocity=(0.0, 0.0, step / SWEEP_DT))
    b.end_world()
    return b.finalize("cpu"), cfg


def test_the_sweep_length_follows_the_thinnest_wall():
    """test_thin_walls.py:37-42: min(SWEEP h, THIN_SWEEP x the thinnest wall)."""
    m, cfg = _channel(3.6e-3)
    thin = _abi.thinnest_wall(m, np.arange(m.shape_count))
    assert thin == pytest.approx(1.0e-3, rel=1e-6)
    assert _abi.sweep_length(cfg.h, thin) == pytest.approx(0.5e-3)
    assert _abi.sweep_length(cfg.h, 1.0) == pytest.approx(_abi.SWEEP * cfg.h)


@pytest.mark.parametrize("step", [3.6e-3, 3.9e-3])
def test_a_step_across_a_1mm_wall_in_a_channel_is_stopped(step):
    """test_thin_walls.py:92-101: steps of 3.6 and 3.9 mm (under SWEEP h = 4 mm, past the wall's far face at 3.5 mm)
    end before the wall's near face at 2.5 mm; a solver that swept only past SWEEP h lets the particle through."""
    m, cfg = _channel(step)
    o = OracleDFSPH(m, cfg, 32)
    o.build_neighbors()
    o.boundary_slots()
    slots = o.buffer(_abi.BUF_SLOT_SHAPE)[:4]
    assert 4 not in slots.tolist() and np.all(slots >= 0)  # the 1 mm wall is not in the full slot row
    o.step(SWEEP_DT)
    assert o.q[0, 2] < 2.5e-3, o.q[0, 2]
