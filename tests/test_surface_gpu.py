"""The plug-in surface around the step (solver.py:343-479): shape-change notification, retiring particles, shape
work and the body-attached patch emitters this library does not provide."""

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


def test_shape_change_rebuilds_like_a_new_solver():
    """After a wall moves, notify_model_changed(SHAPE_PROPERTIES) gives the same steps, bit for bit, as a solver
    created on the changed model with the same state arrays restored. Every tuning option survives the rebuild,
    among them the rule start chosen over an attached learned network (a rebuild that fell back to the default
    learned start would follow another trajectory than the new solver with the same options)."""
    import torch

    from alloy_amd.sph import _abi, scenes
    from alloy_amd.sph.learned import DEFAULT_WEIGHTS, WarmStartMLP
    from alloy_amd.sph.model import ModelFlags

    sc = scenes.dam_break(8, 0.2, 0, device="cuda")
    _log(f"[surface] shape change: {sc.model.particle_count} particles, {sc.model.shape_count} shapes")
    net = WarmStartMLP.load(DEFAULT_WEIGHTS, "cuda")
    opts = {_abi.OPT_CHECK_INTERVAL: 2, _abi.OPT_PAIR_CACHE: _abi.PAIR_CACHE_APPLY, _abi.OPT_LEARNED_SOLVES: 0,
            _abi.OPT_SPARSE_CORRECTION: 0, _abi.OPT_BUILD_BLOCK: 64}
    a = sc.make_solver()
    net.attach(a)
    for o, v in opts.items():
        a.set_option(o, v)
    before = {o: a.get_option(o) for o in _abi.ALL_OPTIONS}
    st_a = sc.model.state()
    for _ in range(5):
        a.step(st_a, st_a, None, None, DT)
    sc.model.shape_transform[0, 0] += 0.01  # move one wall of the tank
    a.notify_model_changed(int(ModelFlags.SHAPE_PROPERTIES))
    after = {o: a.get_option(o) for o in _abi.ALL_OPTIONS}
    assert after == before, (before, after)
    b = sc.make_solver()
    net.attach(b)
    for o, v in opts.items():
        b.set_option(o, v)
    b.alloy_restore_state_arrays(a.alloy_state_arrays(None))
    st_b = sc.model.state()
    st_b.particle_q.copy_(st_a.particle_q)
    st_b.particle_qd.copy_(st_a.particle_qd)
    for _ in range(5):
        a.step(st_a, st_a, None, None, DT)
        b.step(st_b, st_b, None, None, DT)
    torch.cuda.synchronize()
    assert torch.equal(st_a.particle_q, st_b.particle_q) and torch.equal(st_a.particle_qd, st_b.particle_qd)


def test_retire_moves_mass_to_the_ledger():
    import torch

    from alloy_amd.sph import scenes

    sc = scenes.water_column(4, 4, 6, fills=[6, 4], device="cuda")
    s = sc.make_solver()
    st = sc.model.state()
    s.step(st, st, None, None, DT)
    before = s.mass_ledger()
    mask = torch.zeros(sc.model.particle_count, dtype=torch.int32, device="cuda")
    liquid = (sc.model.sph.phase >= 1).nonzero().reshape(-1)
    mask[liquid[:10]] = 1
    s.retire(st, mask)
    after = s.mass_ledger()
    _log(f"[surface] retire: active {before['active']} -> {after['active']}, out {after['out']}")
    assert np.array_equal(after["active"] + after["out"], after["initial"] + after["in"])  # int64, exact
    assert np.array_equal(after["initial"], before["initial"])
    m_fixed = int(round(float(sc.model.particle_mass[0].item()) * 2**40))
    assert after["out"].sum() == 10 * m_fixed
    s.step(st, st, None, None, DT)  # the retired particles are simply absent from the next step
    r = s.report()
    assert np.all(np.isfinite(r.density_err))


def test_shape_work_and_unbuilt_patch_emitters():
    import torch

    from alloy_amd.sph import scenes

    sc = scenes.dam_break(4, 0.1, 0, device="cuda")
    s = sc.make_solver()
    out = torch.full((sc.model.shape_count,), 7.0, dtype=torch.float64, device="cuda")
    s.alloy_shape_work(out)
    assert torch.count_nonzero(out) == 0
    with pytest.raises(NotImplementedError):
        s.add_patch_emitter(None, 0, None)
    with pytest.raises(NotImplementedError):
        s.trigger(0, None)
