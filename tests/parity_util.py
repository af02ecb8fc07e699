"""Helpers for HIP-vs-oracle parity: phase-by-phase runs of SolverDFSPH (GPU) and OracleDFSPH (CPU, float32) on the
same seeded inputs, and field comparisons (bitwise, and norm-wise relative error)."""

from __future__ import annotations

import numpy as np
import torch

from alloy_amd.sph import _abi
from alloy_amd.sph.solver import SolverDFSPH
from oracle.oracle import OracleDFSPH

REL_TOL = 1.0e-5  # per-field norm-wise relative tolerance of one-step parity


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    den = max(float(np.linalg.norm(b)), 1e-30)
    return float(np.linalg.norm(a - b)) / den


def bitwise_fraction(a: np.ndarray, b: np.ndarray) -> float:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.size == 0:
        return 1.0
    return float(np.mean(a.view(np.uint8).reshape(a.size, -1) == b.view(np.uint8).reshape(b.size, -1)))


class Pair:
    """The HIP solver and the float32 oracle on one scene, with host views of both states."""

    def __init__(self, scene, pressure_solver: str | None = None):
        cfg = scene.config
        if pressure_solver is not None:
            import dataclasses

            cfg = dataclasses.replace(cfg, pressure_solver=pressure_solver)
        self.cfg = cfg
        self.model = scene.model
        self.gpu = SolverDFSPH(scene.model, cfg)
        self.orc = OracleDFSPH(scene.model, cfg, 32)
        self.state = scene.model.state()
        self.n = scene.model.particle_count
        self.W = scene.model.world_count

    # ---------------------------------------------------------------- state
    def gq(self):
        torch.cuda.synchronize()
        return self.state.particle_q.cpu().numpy()

    def gqd(self):
        torch.cuda.synchronize()
        return self.state.particle_qd.cpu().numpy()

    def force_oracle_state(self):
        """Teacher forcing: give the oracle the GPU's state and warm starts."""
        self.orc.set_state(self.gq(), self.gqd())
        for which in (_abi.BUF_KV, _abi.BUF_KD):
            self.orc.buffer(which)[:self.n] = self.gbuf(which, np.float32, (self.n,))

    def gbuf(self, which, dtype, shape):
        return self.gpu.buffer(which, dtype, shape)

    def obuf(self, which, count=None):
        b = self.orc.buffer(which)
        return b if count is None else b[:count]

    # ---------------------------------------------------------------- phases
    def call(self, name, *args):
        fn = getattr(self.gpu.lib, name)
        _abi.check(fn(self.gpu._h, *args), self.gpu.lib, name)

    def neighbors(self):
        import ctypes

        self.call("dfsph_build_neighbors", ctypes.c_void_p(self.state.particle_q.data_ptr()), None)
        self.orc.build_neighbors()

    def slots(self):
        import ctypes

        self.call("dfsph_boundary_slots", ctypes.c_void_p(self.state.particle_q.data_ptr()),
                  ctypes.c_void_p(self.state.particle_qd.data_ptr()), None)
        self.orc.boundary_slots()

    def density(self):
        import ctypes

        self.call("dfsph_density_alpha", ctypes.c_void_p(self.state.particle_q.data_ptr()), None)
        self.orc.density_alpha()

    def solve(self, mode, dt, fixed=-1):
        import ctypes

        self.call("dfsph_solve", int(mode), ctypes.c_void_p(self.state.particle_q.data_ptr()),
                  ctypes.c_void_p(self.state.particle_qd.data_ptr()), float(dt), int(fixed), None)
        self.orc.solve(mode, dt, fixed)

    def nonpressure(self, dt):
        import ctypes

        self.call("dfsph_nonpressure", ctypes.c_void_p(self.state.particle_q.data_ptr()),
                  ctypes.c_void_p(self.state.particle_qd.data_ptr()), None, float(dt), None)
        self.orc.nonpressure(dt)

    def advect(self, dt):
        import ctypes

        self.call("dfsph_advect_project", ctypes.c_void_p(self.state.particle_q.data_ptr()),
                  ctypes.c_void_p(self.state.particle_qd.data_ptr()), float(dt), None)
        self.orc.meters()
        self.orc.advect_project(dt)

    def step(self, dt):
        self.gpu.step(self.state, self.state, None, None, dt)
        self.orc.step(dt)

    # ---------------------------------------------------------------- comparisons
    def neighbor_rows(self):
        """(gpu rows, oracle rows) as lists of global ids per particle (valid entries only)."""
        n, cap = self.n, self.cfg.neighbor_capacity
        gc = self.gbuf(_abi.BUF_COUNT, np.int32, (n,))
        oc = self.obuf(_abi.BUF_COUNT, n).copy()
        gn = self.gbuf(_abi.BUF_NBR, np.uint16, (cap, n))
        on = self.obuf(_abi.BUF_NBR, cap * n).reshape(cap, n).copy()
        return gc, oc, gn, on

    def compare_neighbors(self):
        gc, oc, gn, on = self.neighbor_rows()
        assert np.array_equal(gc, oc), f"neighbour counts differ at {np.nonzero(gc != oc)[0][:10]}"
        k = np.arange(gn.shape[0])[:, None]
        valid = k < gc[None, :]
        assert np.array_equal(gn[valid], on[valid]), "neighbour rows differ"
        go = self.gbuf(_abi.BUF_OVERFLOW, np.int32, (self.W,))
        oo = self.obuf(_abi.BUF_OVERFLOW, self.W)
        assert np.array_equal(go, oo), f"overflow {go} vs {oo}"

    def fields(self, spec):
        """spec: [(name, which, comps)] -> {name: (gpu, oracle)}"""
        out = {}
        for name, which, comps in spec:
            shape = (self.n,) if comps == 1 else (self.n, comps)
            g = self.gbuf(which, np.float32, shape)
            o = self.obuf(which, self.n * comps).reshape(shape).copy()
            out[name] = (g, o)
        return out
