"""Autotuner of the native step's launch options on the pack it will run.

Every tuning option (DFSPH_OPT_*: the host convergence-check interval, i.e. the iteration schedule of the pressure
solves, the neighbour-list block size, the pair-gradient cache, the sparse correction pass, the packed acceleration
records, the LDS-staged halos, hipGraph replay, the device-side solve loop) gives
bit-identical results, so the tuner can time them on the live simulation: it snapshots the state, the warm starts,
the ledgers, the emitters and the shape meters, times short step windows for each candidate (one option at a time,
each from the snapshot, median of ``rounds``), keeps the fastest value, and restores the snapshot. The caller's
simulation continues exactly as if no tuning had happened. The reference picks launch parameters inside Warp; this
has no counterpart there.
"""

from __future__ import annotations

import math
import time
from typing import Callable

import numpy as np
import torch

from . import _abi

# (option, candidate values) in the order they are tuned; the first value of each is the library default
CANDIDATES = (
    (_abi.OPT_CHECK_INTERVAL, (1, 2, 3)),
    (_abi.OPT_BUILD_STAGED, (0, 1)),
    (_abi.OPT_BUILD_BLOCK, (128, 64, 256)),
    (_abi.OPT_SPARSE_CORRECTION, (1, 0)),
    (_abi.OPT_PACKED_ACCEL, (1, 0)),
    (_abi.OPT_HALO, (0, 1)),
    (_abi.OPT_PAIR_CACHE, (0, _abi.PAIR_CACHE_APPLY, _abi.PAIR_CACHE_ALL)),
    # graph replay and the device-side solve loop are tuned together: the loop takes the host's per-batch wait off a
    # step, and only then does replaying the step's segments as graphs save the launches' host cost
    ((_abi.OPT_GRAPHS, _abi.OPT_DEVICE_LOOP), ((0, 0), (1, 0), (0, 1), (1, 1))),
)
NAMES = {_abi.OPT_CHECK_INTERVAL: "check_interval", _abi.OPT_BUILD_BLOCK: "build_block",
         _abi.OPT_SPARSE_CORRECTION: "sparse_correction", _abi.OPT_PAIR_CACHE: "pair_cache", _abi.OPT_GRAPHS: "graphs",
         _abi.OPT_LEARNED_SOLVES: "learned_solves", _abi.OPT_PACKED_ACCEL: "packed_accel", _abi.OPT_HALO: "halo",
         _abi.OPT_DEVICE_LOOP: "device_loop", _abi.OPT_BUILD_STAGED: "build_staged",
         _abi.OPT_LEARNED_GATE: "learned_gate"}


class _Snapshot:
    """Everything a step changes: the state, the solver's state arrays (warm starts, flags, phase, emitted and retired
    mass, the emitters' clocks and cursors), the shape meters, the rest of the per-world ledger and the particle
    walls' impulse totals of the current exchange window."""

    def __init__(self, solver, state):
        self.q = state.particle_q.clone()
        self.qd = state.particle_qd.clone()
        self.arrays = solver.alloy_state_arrays(None)
        S, W = solver.model.shape_count, solver.model.world_count
        self.buffers = {_abi.BUF_SHAPE_METER: solver.buffer(_abi.BUF_SHAPE_METER, np.int64, (S, 3))} if S else {}
        for which, dt in ((_abi.BUF_EMIT_OVERFLOW, np.int32), (_abi.BUF_FLUX_E, np.int64), (_abi.BUF_FLUX_G, np.int64)):
            self.buffers[which] = solver.buffer(which, dt, (W,))
        if solver.pw.count:
            self.buffers[_abi.BUF_PW_EXCH] = solver.buffer(_abi.BUF_PW_EXCH, np.int64, (solver.pw.count, 3))

    def restore(self, solver, state):
        state.particle_q.copy_(self.q)
        state.particle_qd.copy_(self.qd)
        solver.alloy_restore_state_arrays(self.arrays)
        for which, a in self.buffers.items():
            solver._write(which, a)


def _window_ms(solver, state, dt: float, warm: int, steps: int) -> float:
    for _ in range(warm):
        solver.step(state, state, None, None, dt)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        solver.step(state, state, None, None, dt)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def autotune(solver, state, dt: float, *, steps: int = 8, warm: int = 2, rounds: int = 2, candidates=CANDIDATES,
             learned: bool = False, margin: float = 0.005, log: Callable[[str], None] | None = None) -> dict:
    """Tune ``solver``'s options on ``state`` in place (the state and the solver come back unchanged) and return
    {"options": {name: value}, "ms_per_step": {name: {value: median ms}}, "seconds": tuning time}.

    ``learned``: with a learned warm start attached, also time it against the reference's rule (DFSPH_OPT_LEARNED_SOLVES
    0). That choice is not bit-identical: both starts converge to the same tolerance, but the iterates differ, so the
    simulation after tuning follows the chosen start (it is restored to the snapshot either way).

    ``margin``: an option keeps its current value unless another is faster by more than this fraction, so that
    timing noise between near-equal values does not move the choice from run to run."""
    t0 = time.perf_counter()
    if learned and getattr(solver, "_warm_model", None) is not None:
        cur = solver.get_option(_abi.OPT_LEARNED_SOLVES)
        if cur:
            candidates = tuple(candidates) + ((_abi.OPT_LEARNED_SOLVES, (cur, 0)),)
    snap = _Snapshot(solver, state)
    # an entry tunes one option (opt, values) or several together ((opt, ...), ((value, ...), ...))
    group = [(tuple(o) if isinstance(o, tuple) else (o,), [tuple(v) if isinstance(v, tuple) else (v,) for v in vals])
             for o, vals in candidates]
    current = {opt: solver.get_option(opt) for opts, _ in group for opt in opts}
    table: dict[str, dict] = {}
    for opts, values in group:
        incumbent = tuple(current[o] for o in opts)
        best_v, best_t, row, t_inc = incumbent, math.inf, {}, math.inf
        label = "+".join(NAMES.get(o, str(o)) for o in opts)
        for v in values:
            trial = {**current, **dict(zip(opts, v))}
            times = []
            for _ in range(rounds):
                snap.restore(solver, state)
                try:
                    for o, val in trial.items():
                        solver.set_option(o, val)
                except ValueError:  # not applicable to this pack (e.g. 256-thread builds with 24-bit rows)
                    times = []
                    break
                times.append(_window_ms(solver, state, dt, warm, steps))
            if not times:
                continue
            key = v[0] if len(v) == 1 else ",".join(str(x) for x in v)
            row[key] = float(np.median(times))
            if v == incumbent:
                t_inc = row[key]
            if row[key] < best_t:
                best_v, best_t = v, row[key]
        if best_v != incumbent and t_inc < math.inf and best_t > t_inc * (1.0 - margin):
            best_v = incumbent  # within the noise of the current value
        current.update(zip(opts, best_v))
        table[label] = row
        if log:
            log(f"autotune {label}: " + ", ".join(f"{k} {t:.3f} ms" for k, t in row.items())
                + f" -> {best_v if len(best_v) > 1 else best_v[0]}")
    for o, val in current.items():
        solver.set_option(o, val)
    snap.restore(solver, state)
    return {"options": {NAMES.get(o, str(o)): v for o, v in current.items()}, "ms_per_step": table,
            "seconds": time.perf_counter() - t0}
