"""Coupling hooks of SolverDFSPH (coupling_hooks.py:37-142): a restart pass starts from the exchange's saved solver
state, so it repeats the first pass bit for bit (warm starts, shape meters, emitters, ledgers, the particle walls'
impulse window), the harvest returns the window's force, and the proxy forms this library does not serve are refused
loudly."""

from __future__ import annotations

import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DT = 1.0e-3
PQ, PQD = 1 << 4, 1 << 5


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pass(sc, s, st, q0, v0, k0, steps, restart):
    """One coupling pass: the coupler sets the state, notifies, and the solver steps."""
    import torch

    st.particle_q.copy_(q0)
    st.particle_qd.copy_(v0)
    s.coupling_notify_input_state_update(st, PQ | PQD, iteration_restart=restart, dt=DT)
    for k in range(steps):
        sc.drive(st, (k0 + k) * DT)
        s.step(st, st, None, None, DT)
    torch.cuda.synchronize()
    out = {"q": st.particle_q.cpu().numpy().copy(), "qd": st.particle_qd.cpu().numpy().copy()}
    out.update(s.alloy_state_arrays(None))
    out["pw_window"] = s.pw.particle_impulse().copy()
    return out


def test_restart_pass_repeats_the_first_pass_bitwise():
    """Four pour worlds (spout, swaying particle-wall bowl, SDF tank), warmed up for 160 steps (the first layers
    reach the bowl); then an exchange of 20 steps is run twice, the second as a restart pass: every state array and
    the impulse window repeat exactly (the shape meters among them: they count the pass that sticks only)."""
    from alloy_amd.sph import scenes

    sc = scenes.pour_worlds(2048, seed=4, device="cuda", subset=range(0, 2048, 512))
    s = sc.make_solver()
    st = sc.model.state()
    for k in range(160):
        sc.drive(st, k * DT)
        s.step(st, st, None, None, DT)
    q0, v0 = st.particle_q.clone(), st.particle_qd.clone()
    first = _pass(sc, s, st, q0, v0, 160, 20, restart=False)
    again = _pass(sc, s, st, q0, v0, 160, 20, restart=True)
    assert set(first) == set(again)
    for key in first:
        assert np.array_equal(first[key], again[key]), key
    assert np.any(first["pw_window"] != 0.0) and "sph.meter_j" in first
    assert np.any(first["sph.mass_in"] != 0)
    _log(f"[coupling] {len(first)} arrays repeat bitwise over a restart pass")


def test_harvest_returns_the_window_force_and_refuses_what_is_not_served():
    import torch

    from alloy_amd.sph import scenes

    sc = scenes.pour_worlds(2048, seed=4, device="cuda", subset=[0])
    s = sc.make_solver()
    st = sc.model.state()
    n = sc.model.particle_count
    s.coupling_notify_input_state_update(st, PQ | PQD, dt=DT)
    for k in range(160):
        sc.drive(st, k * DT)
        s.step(st, st, None, None, DT)
    l2g = np.full(n, -1, np.int64)
    l2g[s.pw.vid] = np.arange(s.pw.count)
    out = torch.zeros((s.pw.count, 3), dtype=torch.float32, device="cuda")
    s.coupling_harvest_proxy_particle_forces(l2g, out, dt=160 * DT)
    expect = s.pw.particle_impulse() / (160 * DT)
    assert np.any(expect != 0.0) and np.allclose(out.cpu().numpy(), expect.astype(np.float32))
    with pytest.raises(ValueError, match="dt > 0"):
        s.coupling_harvest_proxy_particle_forces(l2g, out, dt=0.0)
    with pytest.raises(ValueError, match="staggered"):
        s.coupling_rewind_proxy_particle(l2g, st, out, None, DT)
    s.coupling_harvest_proxy_wrenches(np.zeros(0, np.int64), None, dt=DT)  # no body proxies: nothing to do
    with pytest.raises(NotImplementedError, match="bodies"):
        s.coupling_harvest_proxy_wrenches(np.zeros(1, np.int64), None, dt=DT)
    assert s.coupling_prepare_proxy_contacts(st, None) is None
    assert s.coupling_supports_full_surface_soft_contacts() and not s.coupling_supports_inertial_property_refresh()


def test_staggered_sync_moves_the_walls_back_once():
    """After a notification with new positions the first step moves the wall particles back by v dt before the grid
    (coupling_hooks.py:170-180): equal to stepping from positions moved back by hand, bit for bit."""
    import torch

    from alloy_amd.sph import scenes

    sc = scenes.pour_worlds(2048, seed=4, device="cuda", subset=[0])
    st = sc.model.state()
    sc.drive(st, 0.0)
    a, b = sc.make_solver(), sc.make_solver()
    sa, sb = sc.model.state(), sc.model.state()
    for x in (sa, sb):
        x.particle_q.copy_(st.particle_q)
        x.particle_qd.copy_(st.particle_qd)
    a.coupling_notify_input_state_update(sa, PQ | PQD, dt=DT)
    vid = torch.as_tensor(b.pw.vid, device="cuda")
    sb.particle_q[vid] = sb.particle_q[vid] + (-DT) * sb.particle_qd[vid]
    for s, x in ((a, sa), (b, sb)):
        s.step(x, x, None, None, DT)
        s.step(x, x, None, None, DT)
    torch.cuda.synchronize()
    assert torch.equal(sa.particle_q, sb.particle_q) and torch.equal(sa.particle_qd, sb.particle_qd)
