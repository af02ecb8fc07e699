"""HIP path vs the float32 CPU oracle on identical seeded inputs (the parity tests proper).

Every phase of the step is compared on the same inputs (teacher forcing: before a phase the oracle is given the
GPU's state), so a difference names the phase it comes from. Integer data (neighbour rows, counts, slot shapes,
iteration counts, overflow) must be identical; float fields must agree to 1e-5 norm-wise relative error (written
per assertion). Both sides evaluate the same float32 operation sequence with FMA contraction off, so most fields are
expected bit for bit; the bitwise fraction is printed alongside.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DT = 1.0e-3
TOL = 1.0e-5


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scene(name):
    from alloy_amd.sph import scenes

    if name == "dam_break_8k":
        return scenes.dam_break(20, 0.4, 0, device="cuda")
    if name == "columns_2w":
        return scenes.water_column(6, 6, 12, fills=[12, 7], device="cuda")
    if name == "bowls_3w":
        return scenes.bowl_pool(3, bowl_radius=0.05, side=0.2, device="cuda")
    raise KeyError(name)


def _cmp(pair, spec, what, tol=TOL):
    from tests.parity_util import bitwise_fraction, rel_err

    for name, (g, o) in pair.fields(spec).items():
        e = rel_err(g, o)
        bf = bitwise_fraction(g, o)
        _log(f"  {what}.{name}: rel {e:.3e} bitwise {bf:.4f}")
        assert e <= tol, f"{what}.{name}: relative error {e:.3e} > {tol}"


def _slot_cmp(pair, what):
    from alloy_amd.sph import _abi
    from tests.parity_util import bitwise_fraction, rel_err

    n = pair.n
    gs = pair.gbuf(_abi.BUF_SLOT_SHAPE, np.int32, (n, 4))
    os_ = pair.obuf(_abi.BUF_SLOT_SHAPE, 4 * n).reshape(n, 4)
    assert np.array_equal(gs, os_), f"{what}: slot shapes differ at {np.argwhere(gs != os_)[:5]}"
    valid = gs >= 0
    _log(f"  {what}: {int(valid.sum())} filled slots")
    for name, which, comps in (("psi", _abi.BUF_SLOT_PSI, 1), ("phi", _abi.BUF_SLOT_PHI, 1),
                               ("gpsi", _abi.BUF_SLOT_GPSI, 3), ("xb", _abi.BUF_SLOT_XB, 3),
                               ("vb", _abi.BUF_SLOT_VB, 3), ("nbl", _abi.BUF_SLOT_NBL, 3),
                               ("xbl", _abi.BUF_SLOT_XBL, 3)):
        shape = (n, 4) if comps == 1 else (n, 4, 3)
        g = pair.gbuf(which, np.float32, shape)[valid]
        o = pair.obuf(which, n * 4 * comps).reshape(shape)[valid]
        e = rel_err(g, o) if g.size else 0.0
        _log(f"  {what}.{name}: rel {e:.3e} bitwise {bitwise_fraction(g, o):.4f}")
        assert e <= TOL, f"{what}.{name}: relative error {e:.3e}"


def _world_ints(pair, which, what):
    g = pair.gbuf(which, np.int32, (pair.W,))
    o = pair.obuf(which, pair.W)
    _log(f"  {what}: gpu {g.tolist()[:8]} oracle {o.tolist()[:8]}")
    assert np.array_equal(g, o), f"{what} differ: {g} vs {o}"


@pytest.mark.parametrize("scene_name,solver", [("dam_break_8k", "cg"), ("columns_2w", "cg"),
                                               ("columns_2w", "jacobi"), ("bowls_3w", "cg")])
def test_phase_parity(scene_name, solver):
    from alloy_amd.sph import _abi
    from tests.parity_util import Pair

    _log(f"[phase parity] scene={scene_name} solver={solver}")
    pair = Pair(_scene(scene_name), pressure_solver=solver)
    _log(f"  n={pair.n} W={pair.W}")
    for step in range(3):
        pair.force_oracle_state()
        pair.neighbors()
        pair.compare_neighbors()
        pair.slots()
        _slot_cmp(pair, f"step{step}.slots")
        pair.density()
        _cmp(pair, [("ratio", _abi.BUF_RATIO, 1), ("alpha", _abi.BUF_ALPHA, 1), ("neff", _abi.BUF_NEFF, 1),
                    ("reg", _abi.BUF_REG, 1)], f"step{step}.density")
        g = pair.gbuf(_abi.BUF_N_ACTIVE, np.float64, (pair.W,))
        assert np.array_equal(g, pair.obuf(_abi.BUF_N_ACTIVE, pair.W)), "n_active"
        pair.solve(0, DT)
        _world_ints(pair, _abi.BUF_ITERS_DIV, f"step{step}.iters_div")
        _cmp(pair, [("kappa", _abi.BUF_KAPPA, 1), ("kv", _abi.BUF_KV, 1)], f"step{step}.div")
        from tests.parity_util import rel_err

        e = rel_err(pair.gqd(), pair.orc.qd)
        _log(f"  step{step}.div.qd rel {e:.3e}")
        assert e <= TOL
        pair.force_oracle_state()
        pair.nonpressure(DT)
        e = rel_err(pair.gqd(), pair.orc.qd)
        _log(f"  step{step}.nonpressure.qd rel {e:.3e}")
        assert e <= TOL
        pair.force_oracle_state()
        pair.solve(1, DT)
        _world_ints(pair, _abi.BUF_ITERS_DEN, f"step{step}.iters_den")
        _cmp(pair, [("kappa", _abi.BUF_KAPPA, 1), ("kd", _abi.BUF_KD, 1), ("err", _abi.BUF_ERR, 1)],
             f"step{step}.den")
        e = rel_err(pair.gqd(), pair.orc.qd)
        _log(f"  step{step}.den.qd rel {e:.3e}")
        assert e <= TOL
        pair.force_oracle_state()
        pair.advect(DT)
        eq, ev = rel_err(pair.gq(), pair.orc.q), rel_err(pair.gqd(), pair.orc.qd)
        _log(f"  step{step}.advect q rel {eq:.3e} qd rel {ev:.3e}")
        assert eq <= TOL and ev <= TOL
        _world_ints(pair, _abi.BUF_OVERFLOW, f"step{step}.overflow")


@pytest.mark.parametrize("scene_name", ["dam_break_8k", "bowls_3w"])
def test_step_parity_teacher_forced(scene_name):
    """Whole steps (dfsph_step) from identical states: q, qd, warm starts and per-world reports agree each step."""
    from alloy_amd.sph import _abi
    from tests.parity_util import Pair, bitwise_fraction, rel_err

    _log(f"[step parity] scene={scene_name}")
    pair = Pair(_scene(scene_name))
    for step in range(10):
        pair.force_oracle_state()
        pair.step(DT)
        eq, ev = rel_err(pair.gq(), pair.orc.q), rel_err(pair.gqd(), pair.orc.qd)
        bq = bitwise_fraction(pair.gq(), pair.orc.q)
        _log(f"  step {step}: q rel {eq:.3e} (bitwise {bq:.4f}) qd rel {ev:.3e}")
        assert eq <= TOL and ev <= TOL
        for which in (_abi.BUF_ITERS_DEN, _abi.BUF_ITERS_DIV):
            _world_ints(pair, which, f"step{step}.iters")
        for which in (_abi.BUF_KV, _abi.BUF_KD):
            g = pair.gbuf(which, np.float32, (pair.n,))
            assert rel_err(g, pair.obuf(which, pair.n)) <= TOL


def test_meters_and_report_match():
    """Shape impulse meters (int64 fixed point) and the per-world report after free-running steps."""
    from alloy_amd.sph import _abi
    from tests.parity_util import Pair

    pair = Pair(_scene("columns_2w"))
    for _ in range(5):
        pair.force_oracle_state()
        pair.step(DT)
    S = pair.model.shape_count
    g = pair.gbuf(_abi.BUF_SHAPE_METER, np.int64, (S, 3))
    o = pair.obuf(_abi.BUF_SHAPE_METER, 3 * S).reshape(S, 3)
    rel = np.abs(g - o).max() / max(np.abs(o).max(), 1)
    _log(f"  meters max rel {rel:.3e}")
    assert rel <= 1e-5
    rep = pair.gpu.report()
    assert rep.density_err.shape == (pair.W,)
    assert np.all(rep.density_err <= pair.cfg.max_density_error * (1 + 1e-6) + 1e-9) or np.all(
        rep.iters_density == pair.cfg.max_iterations)


def test_graph_replay_equals_eager_bitwise():
    """dfsph_step's captured hipGraph segments give bitwise the eager launches' results, replay after replay, with
    eager work and syncs between the replays (two solvers on the same scene)."""
    import torch

    from alloy_amd.sph import scenes

    sc = scenes.dam_break(12, 0.3, 0, device="cuda")
    a, b = sc.make_solver(), sc.make_solver()
    a.set_graphs(True)
    b.set_graphs(False)
    sa, sb = sc.model.state(), sc.model.state()
    for k in range(12):
        a.step(sa, sa, None, None, DT)
        b.step(sb, sb, None, None, DT)
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        assert torch.equal(sa.particle_q, sb.particle_q), f"q differs at step {k}"
        assert torch.equal(sa.particle_qd, sb.particle_qd), f"qd differs at step {k}"
        ra, rb = a.report(), b.report()
        assert np.array_equal(ra.iters_density, rb.iters_density)
    _log("  graph replay == eager over 12 steps")
