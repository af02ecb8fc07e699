"""The committed trained warm-start network (alloy_amd/sph/data, trained by alloy_amd.sph.train_warm_start on dam
breaks with randomised block shapes) against the reference warm-start rule on a held-out scene: both reach the same
tolerance every step, and the network needs fewer pressure iterations."""

from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DT = 1.0e-3


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(scene, steps, net=None, solves=3):
    from alloy_amd.sph import _abi

    s = scene.make_solver()
    if net is not None:
        net.attach(s)
        s.set_option(_abi.OPT_LEARNED_SOLVES, solves)
    st = scene.model.state()
    its, den, tol_ok = [], [], True
    cfg = scene.config
    for _ in range(steps):
        s.step(st, st, None, None, DT)
        r = s.report()
        its.append(float(np.mean(r.iters_density) + np.mean(r.iters_divergence)))
        den.append(float(np.mean(r.density_err)))
        capped = r.iters_density >= cfg.max_iterations
        tol_ok &= bool(np.all((r.density_err <= cfg.max_density_error * (1 + 1e-6)) | capped))
    return float(np.mean(its)), float(np.mean(den)), tol_ok


def test_trained_network_lowers_iterations_at_the_same_tolerance():
    from alloy_amd.sph import scenes
    from alloy_amd.sph.learned import DEFAULT_WEIGHTS, WarmStartMLP

    assert os.path.exists(DEFAULT_WEIGHTS), "trained weights are part of the package"
    rec = json.load(open(os.path.splitext(DEFAULT_WEIGHTS)[0] + ".json"))
    net = WarmStartMLP.load(DEFAULT_WEIGHTS, "cuda")
    sc = scenes.dam_break(50, 1.0, 1, device="cuda")  # the 125,000-particle dam break: not in the training data
    plain_it, plain_den, plain_ok = _run(sc, 150)
    learn_it, learn_den, learn_ok = _run(sc, 150, net, rec["learned_solves"])
    _log(f"[learned] iterations per step: plain {plain_it:.3f}, learned {learn_it:.3f}; density error "
         f"{plain_den:.2e} / {learn_den:.2e}")
    assert plain_ok and learn_ok
    assert learn_it < plain_it - 0.3


def test_untrained_network_reproduces_the_rule():
    """A network whose output layer is zero predicts no correction: the learned start equals the reference rule up
    to the sinh/asinh round trip, so the iteration counts match the plain solver's."""
    import torch

    from alloy_amd.sph import scenes
    from alloy_amd.sph.learned import WarmStartMLP
    from oracle import mlp_oracle as mo

    net = WarmStartMLP("cuda", seed=0)
    with torch.no_grad():
        net.params[mo.OFF_W4:] = 0.0
    sc = scenes.dam_break(20, 0.4, 0, device="cuda")
    plain_it, _, _ = _run(sc, 60)
    zero_it, _, ok = _run(sc, 60, net, 3)
    _log(f"[learned] zero network: {zero_it:.3f} iterations per step, plain {plain_it:.3f}")
    assert ok and abs(zero_it - plain_it) < 0.1


def _column_and_dam_break():
    """Two worlds: a short resting column, whose solves sit at the min_iterations floor, and a 20^3 dam break (given a
    downward start in the test, so that its density solve iterates above the floor)."""
    from alloy_amd.sph import scenes
    from alloy_amd.sph.build import add_box_tank, add_fluid_block
    from alloy_amd.sph.model import ModelBuilder

    cfg = scenes.bench_config()
    s = cfg.spacing
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    b.begin_world()
    add_box_tank(b, (0.0, 0.0, 0.0), (6 * s, 6 * s, 16 * s), wall=cfg.h)
    add_fluid_block(b, cfg, (0.0, 0.0, 0.0), (6, 6, 4))
    b.end_world()
    b.begin_world()
    add_box_tank(b, (0.0, 0.0, 0.0), (0.4, 0.4, 0.4), wall=cfg.h)
    add_fluid_block(b, cfg, (0.0, 0.0, 0.0), (20, 20, 20), jitter=1.0e-4, rng=np.random.default_rng(5))
    b.end_world()
    return scenes.Scene(b.finalize("cuda"), cfg)


def test_learned_gate_keeps_worlds_at_the_floor_on_the_rule():
    """With the per-world gate (the default), a world whose fed solve needed no more than min_iterations at the last
    step starts from the reference's rule: the resting column steps bit for bit as with the rule alone, while the dam
    break takes the network's start and both keep the tolerance."""
    import torch

    from alloy_amd.sph import _abi
    from alloy_amd.sph.learned import DEFAULT_WEIGHTS, WarmStartMLP

    rec = json.load(open(os.path.splitext(DEFAULT_WEIGHTS)[0] + ".json"))
    net = WarmStartMLP.load(DEFAULT_WEIGHTS, "cuda")
    sc = _column_and_dam_break()
    rule, learn = sc.make_solver(), sc.make_solver()
    net.attach(learn)
    learn.set_option(_abi.OPT_LEARNED_SOLVES, rec["learned_solves"])
    assert learn.get_option(_abi.OPT_LEARNED_GATE) == 1
    sr, sl = sc.model.state(), sc.model.state()
    ws = sc.model.world_start_np
    for st in (sr, sl):  # the dam break's block starts moving into the floor at 1 m/s: a compressing impact
        st.particle_qd[ws[1]:ws[2], 2] = -1.0
    cfg = sc.config
    differs, gated_on = False, 0
    for k in range(60):
        rule.step(sr, sr, None, None, DT)
        learn.step(sl, sl, None, None, DT)
        rr, rl = rule.report(), learn.report()
        assert rr.iters_density[0] <= cfg.min_iterations and rl.iters_density[0] <= cfg.min_iterations
        assert np.all(rl.density_err <= cfg.max_density_error * (1 + 1e-6))
        torch.cuda.synchronize()
        a, b = slice(ws[0], ws[1]), slice(ws[1], ws[2])
        assert torch.equal(sr.particle_q[a], sl.particle_q[a]) and torch.equal(sr.particle_qd[a], sl.particle_qd[a]), k
        differs |= not torch.equal(sr.particle_q[b], sl.particle_q[b])
        gated_on += int(rl.iters_density[1] > cfg.min_iterations and k < 59)  # the next step takes the network
    _log(f"[learned gate] column world identical to the rule over 60 steps; the dam break took the network's start "
         f"{gated_on} times and differs: {differs}")
    assert gated_on > 0 and differs
