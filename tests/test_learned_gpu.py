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
