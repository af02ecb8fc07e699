"""The step's optional stages together on the GPU, against the oracle bit for bit: Verlet-skin lists, CSF surface
tension with a contact angle on the tank floor, the implicit viscosity with its wall term, a box emitter and a sink,
in two worlds of different liquids, with and without the tuning options."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scene():
    from alloy_amd.sph import scenes
    from alloy_amd.sph.build import add_box_tank, add_fluid_block, add_spares
    from alloy_amd.sph.config import DFSPHConfig
    from alloy_amd.sph.model import ModelBuilder

    cfg = DFSPHConfig(particle_radius=2.5e-4, viscosity="weiler2018", boundary_viscosity=0.5, surface_tension="csf",
                      reneighbor_skin=0.1)
    cfg.validate()
    s = cfg.spacing
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    for k, (mu, sigma, nz) in enumerate(((0.5, 0.06, 8), (0.05, 0.072, 6))):
        b.begin_world()
        tank = add_box_tank(b, (0.0, 0.0, 0.0), (8 * s, 8 * s, 20 * s), wall=cfg.h)
        b._shapes[tank["floor"]]["contact_angle"] = math.radians(70.0 + 30.0 * k)
        add_fluid_block(b, cfg, (s, s, 0.0), (5, 5, nz), viscosity=mu, jitter=0.05 * s,
                        rng=np.random.default_rng(k))
        add_spares(b, cfg, 60, viscosity=mu, park=(4 * s, 4 * s, 18 * s))
        b.end_world()
    m = b.finalize("cuda")
    m.sph.surface_tension.copy_(m.sph.surface_tension.new_tensor(
        np.where(m.sph.phase.cpu().numpy() >= 0, np.where(m.particle_world.cpu().numpy() == 0, 0.06, 0.072), 0.0)))
    sc = scenes.Scene(m, cfg, emitters=[(w, (4 * s, 4 * s, 16 * s, 0.0, 0.0, 0.0, 1.0), (1.0 * s, 1.0 * s, 0.4 * s),
                                         (2.0 * s) ** 2 * 0.5, (0.0, 0.0, -0.5)) for w in range(2)])
    return sc


@pytest.mark.parametrize("options", [{}, {"graphs": 1, "check_interval": 2, "pair_cache": 7, "sparse_correction": 1,
                                         "packed_accel": 0}])
def test_all_stages_match_the_oracle_bitwise(options):
    import torch

    from alloy_amd.sph import _abi
    from tests.test_emitters_gpu import _oracle_for

    sc = _scene()
    s = sc.make_solver()
    for name, val in options.items():
        s.set_option(getattr(_abi, "OPT_" + name.upper()), val)
    sink = s.add_sink(1, (4 * sc.config.spacing, 4 * sc.config.spacing, 0.5 * sc.config.spacing, 0.0, 0.0, 0.0, 1.0),
                      (1.5 * sc.config.spacing, 1.5 * sc.config.spacing, 1.0 * sc.config.spacing))
    assert sink >= 0
    o = _oracle_for(s, sc)
    st = sc.model.state()
    n, W = sc.model.particle_count, sc.model.world_count
    dt = min(5.0e-4, s.config.dt_max)
    rebuilds = 0
    for k in range(40):
        s.step(st, st, None, None, dt)
        o.step(dt)
        torch.cuda.synchronize()
        what = f"step {k}"
        gq, gv = st.particle_q.cpu().numpy(), st.particle_qd.cpu().numpy()
        assert np.array_equal(gq, o.q), f"{what}: positions differ at {np.nonzero((gq != o.q).any(1))[0][:8]}"
        assert np.array_equal(gv, o.qd), f"{what}: velocities differ at {np.nonzero((gv != o.qd).any(1))[0][:8]}"
        assert np.array_equal(sc.model.sph.phase.cpu().numpy(), o.buffer(_abi.BUF_PHASE)[:n]), what
        for which, dtp in ((_abi.BUF_VISC_ITERS, np.int32), (_abi.BUF_ITERS_DEN, np.int32),
                           (_abi.BUF_MASS_IN, np.int64), (_abi.BUF_MASS_OUT, np.int64)):
            assert np.array_equal(s.buffer(which, dtp, (W,)), o.buffer(which)[:W]), f"{what}: buffer {which}"
        assert np.array_equal(s.buffer(_abi.BUF_CSF_KAPPA, np.float32, (n,)), o.buffer(_abi.BUF_CSF_KAPPA)[:n]), what
        rebuilds += int(s.buffer(_abi.BUF_SKIN_FLAG, np.int32, (W,)).sum())
    led = s.mass_ledger()
    assert np.array_equal(led["active"] + led["out"], led["initial"] + led["in"])
    assert np.all(led["in"] > 0)
    _log(f"[combined] {rebuilds} world rebuilds in 40 steps; emitted {led['in'].tolist()}, sunk {led['out'].tolist()}")
    assert rebuilds < 2 * 40


def test_tension_and_wall_viscosity_in_particle_wall_bowls_match_the_oracle_bitwise():
    """Two particle-wall bowls (r = 2.5 mm, small radii) of a viscous liquid with CSF tension and the wall term: the
    colour-field slots feed the curvature's wall turn (free there: particle walls carry no contact angle), the wall
    term and the reaction split; 30 steps bitwise, the walls' exchange impulses too."""
    import dataclasses

    import torch

    from alloy_amd.sph import _abi, scenes
    from alloy_amd.sph.config import DFSPHConfig
    from oracle.oracle import OracleDFSPH

    cfg = DFSPHConfig(particle_radius=2.5e-3, viscosity="weiler2018", boundary_viscosity=0.2, surface_tension="csf",
                      rest_jitter_speed=None)
    base = scenes.bowl_pool(2, bowl_radius=0.03, side=0.1, cfg=cfg, device="cuda", seed=3)
    base.model.sph.viscosity.fill_(0.2)
    liquid = base.model.sph.phase >= 0
    base.model.sph.surface_tension.copy_(torch.where(liquid, 0.07, 0.0).to(torch.float32))
    sc = dataclasses.replace(base, config=cfg)
    s = sc.make_solver()
    o = OracleDFSPH(sc.model, cfg, 32)
    st = sc.model.state()
    n = sc.model.particle_count
    dt = min(1.0e-3, s.config.dt_max)
    for k in range(30):
        s.step(st, st, None, None, dt)
        o.step(dt)
        torch.cuda.synchronize()
        what = f"bowls step {k}"
        gq, gv = st.particle_q.cpu().numpy(), st.particle_qd.cpu().numpy()
        assert np.array_equal(gq, o.q), f"{what}: positions differ at {np.nonzero((gq != o.q).any(1))[0][:8]}"
        assert np.array_equal(gv, o.qd), f"{what}: velocities differ at {np.nonzero((gv != o.qd).any(1))[0][:8]}"
        assert np.array_equal(s.buffer(_abi.BUF_CSF_SURF, np.int32, (n,)), o.buffer(_abi.BUF_CSF_SURF)[:n]), what
    assert s.pw.count > 0
    assert np.array_equal(s.buffer(_abi.BUF_PW_EXCH, np.int64, (s.pw.count * 3,)),
                          o.buffer(_abi.BUF_PW_EXCH)[:s.pw.count * 3])
    assert np.any(s.buffer(_abi.BUF_PW_EXCH, np.int64, (s.pw.count * 3,)) != 0)
