"""Tuning options and profiling of the HIP step: every option value gives bit-identical results (the autotuner may
pick any of them), and the algorithmic bytes the gather kernels count while profiling equal the host formula."""

from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DT = 1.0e-3


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(scene, steps, options):
    import torch

    s = scene.make_solver()
    for opt, val in options.items():
        s.set_option(opt, val)
    st = scene.model.state()
    its = []
    for _ in range(steps):
        s.step(st, st, None, None, DT)
        r = s.report()
        its.append((r.iters_density.copy(), r.iters_divergence.copy()))
    torch.cuda.synchronize()
    return st.particle_q.cpu().numpy(), st.particle_qd.cpu().numpy(), its


@pytest.mark.parametrize("scene_name", ["bowls_3w", "columns_2w"])
def test_options_are_bit_identical(scene_name):
    from alloy_amd.sph import _abi, scenes

    if scene_name == "bowls_3w":
        sc = scenes.bowl_pool(3, bowl_radius=0.05, side=0.2, device="cuda")
    else:
        sc = scenes.water_column(6, 6, 12, fills=[12, 7], device="cuda")
    _log(f"[options] {scene_name}: {sc.model.particle_count} particles, {sc.model.world_count} worlds")
    q0, v0, it0 = _run(sc, 12, {})
    for opts in ({_abi.OPT_PAIR_CACHE: _abi.PAIR_CACHE_ALL}, {_abi.OPT_PAIR_CACHE: _abi.PAIR_CACHE_APPLY},
                 {_abi.OPT_CHECK_INTERVAL: 3}, {_abi.OPT_GRAPHS: 1, _abi.OPT_PAIR_CACHE: _abi.PAIR_CACHE_ACCEL},
                 {_abi.OPT_SPARSE_CORRECTION: 0}, {_abi.OPT_SPARSE_CORRECTION: 0, _abi.OPT_GRAPHS: 1},
                 {_abi.OPT_BUILD_BLOCK: 64}, {_abi.OPT_BUILD_BLOCK: 256}, {_abi.OPT_PACKED_ACCEL: 0},
                 {_abi.OPT_BUILD_STAGED: 1}, {_abi.OPT_BUILD_STAGED: 1, _abi.OPT_BUILD_BLOCK: 64},
                 {_abi.OPT_PACKED_ACCEL: 0, _abi.OPT_GRAPHS: 1, _abi.OPT_PAIR_CACHE: _abi.PAIR_CACHE_ALL},
                 {_abi.OPT_HALO: 1}, {_abi.OPT_HALO: 1, _abi.OPT_PACKED_ACCEL: 0},
                 {_abi.OPT_HALO: 1, _abi.OPT_SPARSE_CORRECTION: 0, _abi.OPT_PAIR_CACHE: _abi.PAIR_CACHE_ALL},
                 {_abi.OPT_HALO: 1, _abi.OPT_GRAPHS: 1, _abi.OPT_CHECK_INTERVAL: 2},
                 {_abi.OPT_DEVICE_LOOP: 1}, {_abi.OPT_DEVICE_LOOP: 1, _abi.OPT_GRAPHS: 1},
                 {_abi.OPT_DEVICE_LOOP: 1, _abi.OPT_HALO: 1, _abi.OPT_PAIR_CACHE: _abi.PAIR_CACHE_ALL,
                  _abi.OPT_SPARSE_CORRECTION: 0}):
        q, v, it = _run(sc, 12, opts)
        _log(f"  {opts}: q equal {np.array_equal(q, q0)} qd equal {np.array_equal(v, v0)}")
        assert np.array_equal(q, q0) and np.array_equal(v, v0), opts
        for (a, b), (c, d) in zip(it, it0):
            assert np.array_equal(a, c) and np.array_equal(b, d), opts


def test_option_abi_round_trip_and_errors():
    from alloy_amd.sph import _abi, scenes

    sc = scenes.dam_break(6, 0.2, 0, device="cuda")
    s = sc.make_solver()
    assert s.get_option(_abi.OPT_PAIR_CACHE) in range(8)
    s.set_option(_abi.OPT_PAIR_CACHE, 5)
    assert s.get_option(_abi.OPT_PAIR_CACHE) == 5
    s.set_option(_abi.OPT_CHECK_INTERVAL, 4)
    assert s.get_option(_abi.OPT_CHECK_INTERVAL) == 4
    with pytest.raises(ValueError):
        s.set_option(_abi.OPT_PAIR_CACHE, 8)
    with pytest.raises(ValueError):
        s.set_option(99, 1)
    assert s.get_option(_abi.OPT_SPARSE_CORRECTION) == 1
    s.set_option(_abi.OPT_SPARSE_CORRECTION, 0)
    assert s.get_option(_abi.OPT_SPARSE_CORRECTION) == 0
    assert s.get_option(_abi.OPT_PACKED_ACCEL) == 1  # one mass per world: on by default
    s.set_option(_abi.OPT_PACKED_ACCEL, 0)
    assert s.get_option(_abi.OPT_PACKED_ACCEL) == 0
    s.set_option(_abi.OPT_PACKED_ACCEL, 1)
    assert s.get_option(_abi.OPT_PACKED_ACCEL) == 1


def test_packed_accel_needs_one_mass_per_world():
    """Two rest densities (so two particle masses) in one world: the packed gathers cannot stand in for the masses,
    so the option is off and refuses to turn on; the step still runs (positions and quotients apart)."""
    import numpy as np

    from alloy_amd.sph import _abi, scenes
    from alloy_amd.sph.build import add_fluid_block
    from alloy_amd.sph.model import ModelBuilder

    cfg = scenes.bench_config(pressure_solver="jacobi")
    b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
    b.begin_world()
    add_fluid_block(b, cfg, (0.0, 0.0, 0.0), (4, 4, 3), rest_density=1000.0)
    add_fluid_block(b, cfg, (0.0, 0.0, 3 * cfg.spacing), (4, 4, 3), rest_density=800.0)
    b.end_world()
    m = b.finalize("cuda")
    s = scenes.Scene(m, cfg).make_solver()
    assert s.get_option(_abi.OPT_PACKED_ACCEL) == 0
    with pytest.raises(ValueError, match="one particle mass per world"):
        s.set_option(_abi.OPT_PACKED_ACCEL, 1)
    st = m.state()
    s.step(st, st, None, None, DT)
    assert np.isfinite(st.particle_q.cpu().numpy()).all()


def test_profiled_bytes_match_the_host_formula():
    """Single world: each k_accel launch either processes every particle (counted bytes = the formula over the
    lists and slots) or exits (0), so the counted total is an integer multiple of the per-launch formula, at most
    the number of timed launches."""
    from alloy_amd.sph import _abi, scenes

    sc = scenes.dam_break(20, 0.4, 0, device="cuda")
    s = sc.make_solver()
    s.set_option(_abi.OPT_SPARSE_CORRECTION, 0)  # every launch of a world does full work (the formula's case)
    st = sc.model.state()
    for _ in range(3):
        s.step(st, st, None, None, DT)
    lib, h = s.lib, s._h
    mask = (1 << _abi.KID_ACCEL) | (1 << _abi.KID_APPLY)
    _abi.check(lib.dfsph_set_profiling(h, mask, 1), lib)
    s.step(st, st, None, None, DT)
    res = {}
    for kid in (_abi.KID_ACCEL, _abi.KID_APPLY):
        c, t, b = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
        _abi.check(lib.dfsph_profile_read(h, kid, ctypes.byref(c), ctypes.byref(t), ctypes.byref(b)), lib)
        res[kid] = (c.value, t.value, b.value)
    _abi.check(lib.dfsph_set_profiling(h, 0, 1), lib)
    n = sc.model.particle_count
    cnt = s.buffer(_abi.BUF_COUNT, np.int32, (n,)).astype(np.int64)
    shp = s.buffer(_abi.BUF_SLOT_SHAPE, np.int32, (n, 4))
    m = sc.model
    liquid = (((m.particle_flags.cpu().numpy() & 1) != 0) & (m.sph.phase.cpu().numpy() >= 1))
    per_accel = int(np.sum(np.where(liquid, 69 + 2 * cnt + 12 * (shp >= 0).sum(1), 17)))
    launches, ms, counted = res[_abi.KID_ACCEL]
    _log(f"[profile] k_accel: {launches} launches, {ms:.3f} ms, counted {counted:.0f} B, per launch {per_accel} B")
    assert launches >= 3 and ms > 0.0
    k = counted / per_accel
    assert abs(k - round(k)) < 1e-9 and 1 <= round(k) <= launches
    la, ma, ba = res[_abi.KID_APPLY]
    assert la >= 3 and ma > 0.0 and ba > 0.0


def test_autotune_leaves_the_simulation_unchanged():
    """The autotuner times candidate options on the live state and restores it: stepping on afterwards gives the
    same bits as a solver that was never tuned."""
    import torch

    from alloy_amd.sph import scenes
    from alloy_amd.sph.autotune import autotune

    sc = scenes.bowl_pool(3, bowl_radius=0.05, side=0.2, device="cuda")
    a, b = sc.make_solver(), sc.make_solver()
    sa, sb = sc.model.state(), sc.model.state()
    for _ in range(3):
        a.step(sa, sa, None, None, DT)
        b.step(sb, sb, None, None, DT)
    assert a.pw.count > 0
    exch = a.pw.particle_impulse()
    assert np.any(exch != 0.0)  # the liquid already pushed on the bowl in this exchange window
    res = autotune(a, sa, DT, steps=2, warm=1, rounds=1, log=_log)
    _log(f"[autotune] {res['options']} in {res['seconds']:.1f} s")
    assert np.array_equal(a.pw.particle_impulse(), exch)  # the tuning steps' wall impulses are not kept
    assert set(res["options"]) == {"check_interval", "build_staged", "build_block", "sparse_correction",
                                   "packed_accel", "halo", "pair_cache", "graphs", "device_loop"}
    for _ in range(6):
        a.step(sa, sa, None, None, DT)
        b.step(sb, sb, None, None, DT)
    torch.cuda.synchronize()
    assert torch.equal(sa.particle_q, sb.particle_q) and torch.equal(sa.particle_qd, sb.particle_qd)
    assert np.array_equal(a.pw.particle_impulse(), b.pw.particle_impulse())
