"""Verlet-skin list reuse on the GPU (neighbors_skin.py; tests/solvers/sph/test_neighbors_skin.py:17-144 of the
reference): the rows the solver reads equal a plain build at every step, the rebuild flags are per world, a superset
that overflows becomes that step's plain row and forces the next rebuild, and a whole run with a skin equals the run
without one, bit for bit, with fewer rebuilds."""

from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Pack:
    """Two worlds of jittered lattice points (4^3 and 5^3, contiguous ids), solvers with and without a skin."""

    def __init__(self, rng, cap: int = 64, skin: float = 0.1):
        import torch

        from alloy_amd.sph import scenes
        from alloy_amd.sph.build import add_fluid_points
        from alloy_amd.sph.config import DFSPHConfig
        from alloy_amd.sph.model import ModelBuilder

        self.cfg = DFSPHConfig(particle_radius=0.005, neighbor_capacity=cap)
        sp = self.cfg.spacing
        b = ModelBuilder(gravity=(0.0, 0.0, -9.81))
        for side in (4, 5):
            ax = np.arange(side, dtype=np.float64)
            g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
            b.begin_world()
            add_fluid_points(b, self.cfg, (g + rng.uniform(-0.2, 0.2, g.shape)) * sp)
            b.end_world()
        self.model = b.finalize("cuda")
        import dataclasses

        self.plain = scenes.Scene(self.model, self.cfg).make_solver()
        self.skin = scenes.Scene(self.model, dataclasses.replace(self.cfg, reneighbor_skin=skin)).make_solver()
        self.q = self.model.particle_q.clone()
        self.sp = sp
        self.torch = torch

    def step(self) -> list[bool]:
        """Build both lists from the current positions; assert equal rows, counts and overflow; return the skin
        solver's rebuilt worlds."""
        from alloy_amd.sph import _abi

        n, W, cap = self.model.particle_count, self.model.world_count, self.cfg.neighbor_capacity
        for s in (self.plain, self.skin):
            _abi.check(s.lib.dfsph_build_neighbors(s._h, ctypes.c_void_p(self.q.data_ptr()), None), s.lib)
        self.torch.cuda.synchronize()
        cp = self.plain.buffer(_abi.BUF_COUNT, np.int32, (n,))
        cs = self.skin.buffer(_abi.BUF_COUNT, np.int32, (n,))
        assert np.array_equal(cp, cs)
        rp = self.plain.buffer(_abi.BUF_NBR, np.uint16, (cap, n))
        rs = self.skin.buffer(_abi.BUF_NBR, np.uint16, (cap, n))
        valid = np.arange(cap)[:, None] < cp[None, :]
        assert np.array_equal(rp[valid], rs[valid])
        assert np.array_equal(self.plain.buffer(_abi.BUF_OVERFLOW, np.int32, (W,)),
                              self.skin.buffer(_abi.BUF_OVERFLOW, np.int32, (W,)))
        return self.skin.buffer(_abi.BUF_SKIN_FLAG, np.int32, (W,)).astype(bool).tolist()


def test_skin_rows_equal_plain_rows_and_flags_are_per_world():
    import torch

    rng = np.random.default_rng(0)
    pk = _Pack(rng)
    assert pk.step() == [True, True]  # first build: every world
    for _ in range(3):  # small moves (< skin / 2 in total): reuse
        pk.q += torch.tensor(rng.uniform(-0.01, 0.01, tuple(pk.q.shape)) * pk.sp, dtype=torch.float32, device="cuda")
        assert pk.step() == [False, False]
    pk.q[70] += 0.3 * pk.sp  # one particle of world 1 moves more than skin / 2
    assert pk.step() == [False, True]
    pk.model.particle_flags[5] = 0  # world 0's active set changes (retired)
    assert pk.step() == [True, False]
    pk.model.particle_flags[5] = 1  # and back (poured)
    assert pk.step() == [True, False]
    assert pk.step() == [False, False]


def test_plain_overflow_is_counted_as_a_plain_build():
    pk = _Pack(np.random.default_rng(1), cap=8)  # plain rows overflow (keep the 8 lowest ids)
    assert pk.step() == [True, True]
    assert pk.plain.report().neighbor_overflow.sum() > 0
    assert pk.step() == [False, False]  # the superset fits: reused, overflow counted again


def test_full_superset_rows_are_plain_rows_and_force_the_next_rebuild():
    """Skin = h: within 2 h a particle of the 5^3 world has up to 124 neighbours, more than the 64 a superset row
    holds, so its rows are built plain and the world rebuilds again with no motion; the 4^3 world's supersets (at
    most 63 entries) fit and are reused."""
    pk = _Pack(np.random.default_rng(2), skin=1.0)
    assert pk.step() == [True, True]
    assert pk.step() == [False, True]
    assert pk.step() == [False, True]


def test_particle_walls_refuse_a_skin():
    import dataclasses

    from alloy_amd.sph import scenes

    sc = scenes.bowl_pool(1, device="cuda")
    with pytest.raises(ValueError, match="reneighbor_skin"):
        scenes.Scene(sc.model, dataclasses.replace(sc.config, reneighbor_skin=0.1)).make_solver()


@pytest.mark.parametrize("what", ["dam_break", "column"])
def test_solver_with_a_skin_matches_the_plain_lists_bitwise(what):
    """test_neighbors_skin.py:120-144: a small dam break and a resting column give the same end state, bit for bit,
    with reneighbor_skin 0.1 as with a rebuild every step, and the skin rebuilds on fewer steps."""
    import dataclasses

    import torch

    from alloy_amd.sph import _abi, scenes

    base = scenes.dam_break(8, 0.2, 0, device="cuda") if what == "dam_break" else scenes.water_column(
        4, 4, 8, device="cuda")
    ends, rebuilds = [], 0
    for skin in (None, 0.1):
        s = scenes.Scene(base.model, dataclasses.replace(base.config, reneighbor_skin=skin)).make_solver()
        st = base.model.state()
        s.prime(st, 1.0e-3)
        for _ in range(60):
            s.step(st, st, None, None, 1.0e-3)
            if skin is not None:
                rebuilds += int(s.buffer(_abi.BUF_SKIN_FLAG, np.int32, (1,))[0])
        torch.cuda.synchronize()
        ends.append(st.particle_q.clone())
    _log(f"[skin] {what}: {rebuilds} rebuilds in 60 steps")
    assert torch.equal(ends[0], ends[1])
    assert 0 < rebuilds < 60
