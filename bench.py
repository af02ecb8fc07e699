"""Benchmark of the DFSPH liquid step on one MI355X: particle-steps/s at the 0.1 % density-error tolerance.

Default workload: the 125,000-particle dam break (50^3 lattice block, spacing 0.01 m, jitter U(+-1e-4) m from seed
1, in one corner of a 1.0 m SDF box tank), plain DFSPH, fp32 physics, dt 1 ms, 500 timed steps after 20 warm-up
steps. Inputs are resident in HBM when the timed region starts; the timed region is bracketed by a device
synchronisation and timed with HIP events on the solver's stream. One JSON line is printed.

The line also carries
- "roofline": the dominant kernel's algorithmic bytes per launch / its average launch duration, both measured live
  over the timed region with HIP events around each launch on the launch stream; "traffic" is the HBM bytes per
  launch from rocprofv3 FETCH_SIZE + WRITE_SIZE counters (profiles/), or null when no counter run is recorded;
- "cpu_baseline": the CPU oracle (the same algorithm, float32, one core) timed on this host on a bounded sample of
  the same workload;
- "fp64_density_err": mean max(rho/rho0 - 1, 0) over the liquid of the final state, evaluated by the FP64 oracle.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "dam_break_125k": dict(kind="dam_break", n_side=50, box=1.0, seed=1),
    "dam_break_8k": dict(kind="dam_break", n_side=20, box=0.4, seed=0),
    "dam_break_1m": dict(kind="dam_break", n_side=100, box=2.0, seed=3),
    "bowl_worlds_512": dict(kind="bowl_worlds", worlds=512, seed=2),
}


def make_scene(wl: dict, device: str):
    from alloy_amd.sph import scenes

    if wl["kind"] == "dam_break":
        return scenes.dam_break(wl["n_side"], wl["box"], wl["seed"], device=device)
    return scenes.bowl_worlds(wl["worlds"], seed=wl["seed"], device=device)


def describe(wl: dict, sc) -> str:
    n = sum(sc.fluid_counts)
    if wl["kind"] == "dam_break":
        return (f"plain DFSPH dam break, {n} particles ({wl['n_side']}^3 lattice, spacing 0.01 m, jitter 1e-4 m, seed "
                f"{wl['seed']}) in a {wl['box']} m SDF box tank, dt 1 ms, 0.1 % density and divergence tolerance, "
                "projected CG, 1 world")
    return (f"batched DFSPH, {wl['worlds']} worlds in one pack (box tanks of side U(0.3, 0.6) m, a static particle-wall "
            f"bowl of radius 0.08 m each, U(2000, 20000) liquid particles each, seed {wl['seed']}): {n} liquid "
            f"particles, {sc.model.particle_count - n} wall particles, dt 1 ms, 0.1 % tolerances, projected CG")


def _log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel: str, n_fluid: int, n_other: int, sum_count: int, n_slots: int) -> float:
    """Compulsory bytes of one launch over the step's particles: every per-particle array the kernel reads or
    writes, once; neighbour attributes counted in their owner's read; neighbour entries 2 B (uint16 rows); slot
    gradients 12 B per filled slot, the slot-shape row 16 B per particle; non-liquid particles read their flags and
    phase (and write zeros)."""
    if kernel == "k_apply":
        # q 12, acc 12, mass 4, rho0 4, count 4, flags 4, phase 1, world 4, reg 4, vec 4, out 4, slot shapes 16
        return n_fluid * 77.0 + 2.0 * sum_count + 12.0 * n_slots + n_other * 9.0
    if kernel == "k_accel":
        # q 12, vec 4, rho0 4, mass 4, count 4, flags 4, phase 1, world 4, slot shapes 16, acc out 12
        return n_fluid * 65.0 + 2.0 * sum_count + 12.0 * n_slots + n_other * 17.0
    raise KeyError(kernel)


def cpu_baseline(workload: dict, steps: int, warmup: int) -> dict:
    from alloy_amd.sph import scenes
    from oracle.oracle import OracleDFSPH

    sc = make_scene(workload, "cpu")
    o = OracleDFSPH(sc.model, sc.config, 32)
    for _ in range(warmup):
        o.step(scenes.DT)
    t = time.perf_counter()
    for _ in range(steps):
        o.step(scenes.DT)
    dt = time.perf_counter() - t
    n = sum(sc.fluid_counts)
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle (float32, serial C++ -O3) on the same scene: {n} particles, {warmup} warm-up + "
                      f"{steps} timed steps ({dt:.1f} s)"}


def fp64_density_error(scene, q: np.ndarray, qd: np.ndarray) -> float:
    from alloy_amd.sph import _abi
    from oracle.oracle import OracleDFSPH

    o = OracleDFSPH(scene.model, scene.config, 64)
    o.set_state(q.astype(np.float64), qd.astype(np.float64))
    o.build_neighbors()
    o.boundary_slots()
    o.density_alpha()
    r = o.buffer(_abi.BUF_RATIO)[: o.n]
    a = o.buffer(_abi.BUF_ACTIVE)[: o.n] > 0
    return float(np.mean(np.maximum(r[a] - 1.0, 0.0))) if a.any() else 0.0


def traffic_from_profiles(kernel: str, workload: str):
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return float(json.load(f)["hbm_bytes_per_launch"])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dam_break_125k", choices=sorted(WORKLOADS))
    ap.add_argument("--check-every", type=int, default=1, help="iterations between host convergence checks")
    ap.add_argument("--graphs", action="store_true", help="replay the step's segments as hipGraphs")
    ap.add_argument("--pair-cache", type=int, default=None, help="DFSPH_PAIR_CACHE_* mask (default: the library's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB (3 launches) and exit: the known-size "
                         "kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (worlds shard across "
              "GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        return 2

    import torch

    if args.copy_probe_mib > 0:
        nbytes = args.copy_probe_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "launches": 3}))
        return 0

    from alloy_amd.sph import _abi, scenes

    wl = WORKLOADS[args.workload]
    _log(f"[bench] workload {args.workload} {wl} steps={args.steps} warmup={args.warmup}")
    sc = make_scene(wl, "cuda")
    solver = sc.make_solver()
    solver.set_check_interval(args.check_every)
    solver.set_graphs(args.graphs)
    if args.pair_cache is not None:
        solver.set_option(_abi.OPT_PAIR_CACHE, args.pair_cache)
    state = sc.model.state()
    n_fluid = sum(sc.fluid_counts)
    W = sc.model.world_count
    n_all = sc.model.particle_count
    dt = scenes.DT
    for _ in range(args.warmup):
        solver.step(state, state, None, None, dt)
    torch.cuda.synchronize()
    _log("[bench] warm-up done")
    lib, h = solver.lib, solver._h
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    den, it_den, it_div, unconv = [], [], [], 0
    torch.cuda.synchronize()
    e0.record(stream)
    for k in range(args.steps):
        solver.step(state, state, None, None, dt)
        if k % 10 == 9:
            r = solver.report()
            den.append(float(np.mean(r.density_err)))
            it_den.append(float(np.mean(r.iters_density)))
            it_div.append(float(np.mean(r.iters_divergence)))
            unconv += int(np.sum(r.unconverged))
    e1.record(stream)
    torch.cuda.synchronize()
    total_ms = e0.elapsed_time(e1)
    _log(f"[bench] timed {args.steps} steps in {total_ms:.1f} ms")
    # per-kernel durations: HIP events around every launch of the main kernels on the launch stream, over a second
    # window of the same run (event records need eager launches; the kernels are the same)
    prof_steps = max(min(args.steps, 50), 1)
    _abi.check(lib.dfsph_set_profiling(h, 1), lib)
    for _ in range(prof_steps):
        solver.step(state, state, None, None, dt)
    torch.cuda.synchronize()
    kstats = {}
    for kid, name in enumerate(_abi.KERNEL_NAMES):
        c, t = ctypes.c_int32(), ctypes.c_double()
        _abi.check(lib.dfsph_profile_read(h, kid, ctypes.byref(c), ctypes.byref(t)), lib)
        kstats[name] = (int(c.value), float(t.value))
    _abi.check(lib.dfsph_set_profiling(h, 0), lib)
    # per-launch algorithmic bytes from the final state's lists and slots (K and filled slots drift slowly)
    cnt = solver.buffer(_abi.BUF_COUNT, np.int32, (n_all,))
    shp = solver.buffer(_abi.BUF_SLOT_SHAPE, np.int32, (n_all, 4))
    sum_count, n_slots = int(cnt.sum()), int((shp >= 0).sum())
    dom = max(("k_apply", "k_accel"), key=lambda k: kstats[k][1])
    nl, tl = kstats[dom]
    avg_ms = tl / max(nl, 1)
    bytes_launch = algorithmic_bytes(dom, n_fluid, n_all - n_fluid, sum_count, n_slots)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    value = n_fluid * args.steps / (total_ms * 1e-3)
    steps_total = args.steps + prof_steps
    out = {
        "metric": "particle-steps/s at 0.1% density error, learned+batched vs plain DFSPH, 1 GPU",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": describe(wl, sc), "particles": n_fluid, "worlds": W,
                   "solver": "plain (no learned warm start)"},
        "density_err_mean": float(np.mean(den)) if den else None,
        "density_err_max_sampled": float(np.max(den)) if den else None,
        "iters_density_mean": float(np.mean(it_den)) if it_den else None,
        "iters_divergence_mean": float(np.mean(it_div)) if it_div else None,
        "unconverged_sampled_steps": unconv,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles(dom, args.workload),
                     "alg_bytes_per_launch": bytes_launch, "avg_launch_ms": avg_ms, "launches": nl,
                     "mean_neighbors": sum_count / max(n_fluid, 1), "filled_slots": n_slots},
        "kernel_ms_per_step": {k: round(v[1] / prof_steps, 4) for k, v in kstats.items()},
        "kernel_launches_per_step": {k: round(v[0] / prof_steps, 2) for k, v in kstats.items()},
        "profiled_steps": prof_steps,
    }
    if not args.no_fp64:
        torch.cuda.synchronize()
        out["fp64_density_err"] = fp64_density_error(sc, state.particle_q.cpu().numpy(),
                                                     state.particle_qd.cpu().numpy())
        _log(f"[bench] fp64 density error {out['fp64_density_err']:.3e}")
    if not args.no_cpu_baseline:
        _log("[bench] timing the CPU oracle")
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_steps, 2)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
