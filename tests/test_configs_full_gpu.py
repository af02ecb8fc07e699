"""The benchmark configurations at their full sizes on the GPU, against the oracle:

- the 1,000,000-particle dam break (100^3 lattice, 24-bit rows): neighbour rows bit-exact at steps 0 and 50, one
  teacher-forced whole step bit for bit, and the learned warm start needing fewer pressure iterations than the
  reference's rule at the same tolerance over 120 steps;
- the 2,048-world pour pack (about 7.4 M particles with the spares and the swaying particle-wall bowls): the first 30
  steps bit for bit against the oracle (emission starts at step 5), then 200 steps with the int64 mass ledger exact;
- the 512-world bowl pack (6.2 M particles): one teacher-forced whole step bit for bit after 30 steps.
"""

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


def _teacher_forced_step(s, st, sc, what):
    """Give the oracle the GPU's state and warm starts, step both once, compare positions and velocities bitwise."""
    import torch

    from alloy_amd.sph import _abi
    from oracle.oracle import OracleDFSPH

    torch.cuda.synchronize()
    n = sc.model.particle_count
    _log(f"  {what}: oracle of {n} particles")
    o = OracleDFSPH(sc.model, sc.config, 32)
    o.set_state(st.particle_q.cpu().numpy(), st.particle_qd.cpu().numpy())
    for which in (_abi.BUF_KV, _abi.BUF_KD):
        o.buffer(which)[:n] = s.buffer(which, np.float32, (n,))
    s.step(st, st, None, None, DT)
    torch.cuda.synchronize()
    _log(f"  {what}: GPU step done, oracle step")
    o.step(DT)
    gq, gv = st.particle_q.cpu().numpy(), st.particle_qd.cpu().numpy()
    bad_q = np.nonzero((gq != o.q).any(1))[0]
    bad_v = np.nonzero((gv != o.qd).any(1))[0]
    assert bad_q.size == 0, f"{what}: {bad_q.size} positions differ, first {bad_q[:8]}"
    assert bad_v.size == 0, f"{what}: {bad_v.size} velocities differ, first {bad_v[:8]}"
    W = sc.model.world_count
    for which in (_abi.BUF_ITERS_DEN, _abi.BUF_ITERS_DIV):
        assert np.array_equal(s.buffer(which, np.int32, (W,)), o.buffer(which)[:W]), f"{what}: iteration counts"
    _log(f"  {what}: identical, bit for bit")


def test_dam_break_1m_lists_at_steps_0_and_50():
    from alloy_amd.sph import scenes
    from oracle.oracle import OracleDFSPH
    from tests.test_lists_full_gpu import _compare_lists

    sc = scenes.dam_break(100, 2.0, 3, device="cuda")
    s = sc.make_solver()
    st = sc.model.state()
    o = OracleDFSPH(sc.model, sc.config, 32)
    _log(f"[configs] 1M dam break: {sc.model.particle_count} particles, 24-bit rows")
    _compare_lists(s, o, st, sc.config.neighbor_capacity, True, "step 0")
    for _ in range(50):
        s.step(st, st, None, None, DT)
    _compare_lists(s, o, st, sc.config.neighbor_capacity, True, "step 50")


def test_dam_break_1m_teacher_forced_step():
    from alloy_amd.sph import scenes

    sc = scenes.dam_break(100, 2.0, 3, device="cuda")
    s = sc.make_solver()
    st = sc.model.state()
    _log(f"[configs] 1M dam break, 40 steps then a teacher-forced step")
    for _ in range(40):
        s.step(st, st, None, None, DT)
    _teacher_forced_step(s, st, sc, "1M step 41")


def _iterations(sc, steps, net=None, solves=3):
    from alloy_amd.sph import _abi

    s = sc.make_solver()
    if net is not None:
        net.attach(s)
        s.set_option(_abi.OPT_LEARNED_SOLVES, solves)
    st = sc.model.state()
    cfg = sc.config
    its, ok = [], True
    for _ in range(steps):
        s.step(st, st, None, None, DT)
        r = s.report()
        its.append((int(r.iters_density[0]), int(r.iters_divergence[0])))
        capped = r.iters_density >= cfg.max_iterations
        ok &= bool(np.all((r.density_err <= cfg.max_density_error * (1 + 1e-6)) | capped))
    return np.array(its, np.float64), ok


def test_dam_break_1m_learned_start_lowers_iterations():
    """The committed network (trained on dam breaks of other shapes and on small bowl packs) against the rule on the
    1M dam break, 120 steps each: both meet the density tolerance every step; the learned start needs fewer
    iterations per step, counted over both solves."""
    import json
    import os

    from alloy_amd.sph import scenes
    from alloy_amd.sph.learned import DEFAULT_WEIGHTS, WarmStartMLP

    rec = json.load(open(os.path.splitext(DEFAULT_WEIGHTS)[0] + ".json"))
    net = WarmStartMLP.load(DEFAULT_WEIGHTS, "cuda")
    sc = scenes.dam_break(100, 2.0, 3, device="cuda")
    _log("[configs] 1M dam break: rule start, 120 steps")
    rule, rule_ok = _iterations(sc, 120)
    _log("[configs] 1M dam break: learned start, 120 steps")
    learned, learned_ok = _iterations(sc, 120, net, rec["learned_solves"])
    r_tot, l_tot = rule.sum(1).mean(), learned.sum(1).mean()
    _log(f"[configs] iterations per step (density + divergence): rule {rule.mean(0)} = {r_tot:.3f}, learned "
         f"{learned.mean(0)} = {l_tot:.3f}")
    assert rule_ok and learned_ok
    assert l_tot < r_tot - 0.5


def test_pour_worlds_2048_full_pack():
    import torch

    from alloy_amd.sph import scenes
    from tests.test_emitters_gpu import _compare, _exact, _oracle_for

    sc = scenes.pour_worlds(2048, seed=4, device="cuda")
    s = sc.make_solver()
    st = sc.model.state()
    o = _oracle_for(s, sc)
    _log(f"[configs] pour pack: {sc.model.world_count} worlds, {sc.model.particle_count} particles")
    for k in range(30):
        sc.drive(st, k * DT)
        torch.cuda.synchronize()
        o.qd[...] = st.particle_qd.cpu().numpy()  # the same wall velocities for the oracle
        s.step(st, st, None, None, DT)
        o.step(DT)
        if k % 10 == 9:
            _compare(s, o, st, f"step {k + 1}")
            _log(f"  step {k + 1}: identical, bit for bit")
    del o
    for k in range(30, 200):
        sc.drive(st, k * DT)
        s.step(st, st, None, None, DT)
        if k % 50 == 49:
            assert _exact(s.mass_ledger()), f"ledger at step {k + 1}"
            _log(f"  step {k + 1}: ledger exact")
    r = s.report()
    assert np.all(r.mass_in > 0) and _exact(s.mass_ledger())
    assert np.all(np.isfinite(st.particle_q.cpu().numpy()))


def test_bowl_worlds_512_teacher_forced_step():
    from alloy_amd.sph import scenes

    sc = scenes.bowl_worlds(512, seed=2, device="cuda")
    s = sc.make_solver()
    st = sc.model.state()
    _log(f"[configs] 512 bowl worlds: {sc.model.particle_count} particles, 30 steps then a teacher-forced step")
    for _ in range(30):
        s.step(st, st, None, None, DT)
    _teacher_forced_step(s, st, sc, "bowl pack step 31")
