"""Benchmark of the DFSPH liquid step on one MI355X: particle-steps/s at the 0.1 % density-error tolerance, the
batched learned-warm-start solver against plain single-world DFSPH.

Default workload: 512 batched worlds in one pack, each a box tank of side U(0.3, 0.6) m with a static two-layer
particle-wall bowl (hemisphere, radius 0.08 m) and a seeded lattice block of U{2,000..20,000} liquid particles
(5,672,431 in all, plus 549,888 wall particles), fp32 physics, dt 1 ms, 300 timed steps after 20 warm-up steps,
every world one pressure solve per phase launch. The learned warm start (the committed trained network) feeds the
solves its training run chose (the density solve), gated per world, when its weights are present; the autotuner,
run between the warm-up and the timed region, times it against the reference's rule and keeps the faster.

`value` is the batched (+learned) run's liquid particle-steps per second: inputs resident in HBM, the timed region
bracketed by device synchronisations and timed with HIP events on the solver's stream. The line also carries
- "regime": the steps the timed region covered, the solves' mean iterations and density errors there;
- "plain": the same worlds stepped as plain single-world DFSPH (one solver, i.e. one launch per phase, per world,
  no learned start), on an even sample of the worlds, and `speedup_vs_plain`;
- "batched_rule_start" / "batched_learned_start": the same batched run with the other warm start, so the learned
  start's own effect is on the line; "single_launch_rule_start": one batched launch per phase with the library's
  default options and the rule, and `vs_single_launch`;
- "dam_break_1m_learned_vs_rule": the learned start against the rule on the 1,000,000-particle dam break;
- "roofline": the dominant kernel's algorithmic bytes, counted by the kernels themselves over the window after the
  timed region, divided by its summed launch durations (HIP events inside each launch's dispatch, so they agree with
  rocprofv3's kernel trace); "traffic" is the HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE runs
  (profiles/), or null without one for this run's steps and options;
- "cpu_baseline": the CPU oracle (the same algorithm, float32) on the box's cores (OpenMP) and on one core, on a
  bounded sample of the workload;
- "fp64_density_err": mean max(rho/rho0 - 1, 0) of the final state's liquid, evaluated by the FP64 oracle.
`--workload dam_break_125k` (the 125,000-particle dam break) and the others measure a single world;
`--workload pour_worlds_2048` the batched pour pack (2,048 worlds, each a spout filling a swaying particle-wall bowl
from a pool of U{1,000..4,000} spares, 1,000 steps): there `value` counts the liquid active at each step (particle-
steps/s = sum over steps of the active liquid / seconds), the caller's per-step wall drive inside the timed region.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s at 0.1% density error, learned+batched vs plain DFSPH, 1 GPU"
WORKLOADS = {  # steps: the timed steps each workload is specified with
    "bowl_worlds_512": dict(kind="bowl_worlds", worlds=512, seed=2, steps=300),
    "dam_break_125k": dict(kind="dam_break", n_side=50, box=1.0, seed=1, steps=500),
    "dam_break_8k": dict(kind="dam_break", n_side=20, box=0.4, seed=0, steps=200),
    "dam_break_1m": dict(kind="dam_break", n_side=100, box=2.0, seed=3, steps=200),
    "pour_worlds_2048": dict(kind="pour_worlds", worlds=2048, seed=4, steps=1000),
}
CPU_SAMPLE_PARTICLE_STEPS = 3.0e6  # about 20 s of the serial oracle
OPTION_IDS = {"check_interval": 1, "graphs": 2, "pair_cache": 3, "learned_solves": 4, "sparse_correction": 5,
              "build_block": 6, "packed_accel": 7, "halo": 8, "device_loop": 9, "build_staged": 10,
              "learned_gate": 11}


def make_scene(wl: dict, device: str, subset=None):
    from alloy_amd.sph import scenes

    if wl["kind"] == "dam_break":
        return scenes.dam_break(wl["n_side"], wl["box"], wl["seed"], device=device)
    if wl["kind"] == "pour_worlds":
        return scenes.pour_worlds(wl["worlds"], seed=wl["seed"], device=device, subset=subset)
    return scenes.bowl_worlds(wl["worlds"], seed=wl["seed"], device=device, subset=subset)


class Run:
    """A scene's solver and state plus the caller's per-step work (the swaying bowls of pour worlds)."""

    def __init__(self, sc, solver, state):
        self.sc, self.solver, self.state, self.t = sc, solver, state, 0.0

    def step(self, dt: float) -> None:
        self.sc.drive(self.state, self.t)
        self.solver.step(self.state, self.state, None, None, dt)
        self.t += dt

    def active(self):
        """Active liquid particles now, as a device scalar (no sync)."""
        m = self.solver.model
        return (((m.particle_flags & 1) != 0) & (m.sph.phase >= 1)).sum()


def describe(wl: dict, sc) -> str:
    n = sum(sc.fluid_counts)
    if wl["kind"] == "pour_worlds":
        return (f"{wl['worlds']} batched pour worlds (0.32 m box tanks, a two-layer particle-wall bowl of radius 0.08 m "
                "swaying along x with amplitude 0.05 m and period 1 s at a phase U(0, 2 pi) per world, a spout at "
                "z = 0.25 m with half extents 0.01 m emitting a 2 x 2 layer at 2 m/s every 5 ms, U{1000..4000} spares "
                f"per world, escape box the tank + 0.1 m, seed {wl['seed']}): {n} spares, "
                f"{sc.model.particle_count - n} wall particles, no liquid at step 0, dt 1 ms, 0.1 % tolerances, "
                "projected CG")
    if wl["kind"] == "dam_break":
        return (f"dam break, {n} particles ({wl['n_side']}^3 lattice, spacing 0.01 m, jitter 1e-4 m, seed {wl['seed']}) "
                f"in a {wl['box']} m SDF box tank, dt 1 ms, 0.1 % density and divergence tolerance, projected CG, "
                "1 world")
    return (f"{wl['worlds']} batched worlds (box tanks of side U(0.3, 0.6) m, a static particle-wall bowl of radius "
            f"0.08 m each, U(2000, 20000) liquid particles each, seed {wl['seed']}): {n} liquid particles, "
            f"{sc.model.particle_count - n} wall particles, dt 1 ms, 0.1 % tolerances, projected CG")


def _log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def sample_worlds(W: int, k: int) -> list[int]:
    return sorted(set(np.linspace(0, W - 1, min(k, W)).round().astype(int).tolist()))


def load_learned(mode: str, device: str):
    """The committed trained warm-start network and the solves its training run chose to feed, or (None, 0) (mode
    auto: when its weights exist)."""
    from alloy_amd.sph.learned import DEFAULT_WEIGHTS, WarmStartMLP

    if mode == "off" or (mode == "auto" and not os.path.exists(DEFAULT_WEIGHTS)):
        return None, 0
    rec = os.path.splitext(DEFAULT_WEIGHTS)[0] + ".json"
    solves = json.load(open(rec)).get("learned_solves", 3) if os.path.exists(rec) else 3
    return WarmStartMLP.load(DEFAULT_WEIGHTS, device), int(solves)


def device_view(solver, which: int, typestr: str, count: int):
    """A torch view (no copy) of a library buffer in device memory."""
    import torch

    from alloy_amd.sph import _abi

    ptr, nbytes = ctypes.c_void_p(), ctypes.c_int64()
    _abi.check(solver.lib.dfsph_get_buffer(solver._h, which, ctypes.byref(ptr), ctypes.byref(nbytes)), solver.lib)

    class _Export:
        __cuda_array_interface__ = {"shape": (count,), "typestr": typestr, "data": (ptr.value, False), "version": 3}

    return torch.as_tensor(_Export(), device=solver.device)


def timed_run(run: Run, steps: int, dt: float, prof_mask: int = 0, sample: int = 1, count_active: bool = False) -> dict:
    """`steps` steps timed with HIP events on the current stream, a report (the step's only host sync point) every
    10 steps; over the same region, per-launch timing of a pseudo-random 1 / `sample` of the launches of the kernels
    in `prof_mask` (a timed launch costs the host more than a plain one; sampling keeps that out of the step time).
    Both solves' per-world iteration counts are summed on the device after every step (two small adds, no sync), so
    the iteration means cover every step. `count_active`: sum the active liquid after every step on the device
    (particle-steps of a growing pour)."""
    import torch

    from alloy_amd.sph import _abi

    solver = run.solver
    lib, h = solver.lib, solver._h
    stream = torch.cuda.current_stream()
    if prof_mask:
        _abi.check(lib.dfsph_set_profiling(h, prof_mask, sample), lib)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    den, unconv = [], 0
    W = solver.model.world_count
    its = [device_view(solver, b, "<i4", W) for b in (_abi.BUF_ITERS_DEN, _abi.BUF_ITERS_DIV)]
    it_sum = [torch.zeros(W, dtype=torch.int64, device=solver.device) for _ in its]
    acc = torch.zeros((), dtype=torch.int64, device=solver.device) if count_active else None
    # the bookkeeping's torch kernels run once before the timed region: a kernel's first launch in a process loads its
    # code object, which on a freshly booted machine (cold page cache) takes tens of ms
    for a, v in zip(it_sum, its):
        a += v
        a.zero_()
    if acc is not None:
        acc += run.active()
        acc.zero_()
    torch.cuda.synchronize()
    e0.record(stream)
    for k in range(steps):
        run.step(dt)
        for a, v in zip(it_sum, its):
            a += v
        if acc is not None:
            acc += run.active()
        if k % 10 == 9:
            r = solver.report()
            den.append(float(np.mean(r.density_err)))
            unconv += int(np.sum(r.unconverged))
    e1.record(stream)
    torch.cuda.synchronize()
    it_mean = [float(a.sum().item()) / (W * steps) for a in it_sum]
    out = {"total_ms": e0.elapsed_time(e1), "density_err_mean": float(np.mean(den)) if den else None,
           "density_err_max_sampled": float(np.max(den)) if den else None,
           "iters_density_mean": it_mean[0], "iters_divergence_mean": it_mean[1], "unconverged_sampled_steps": unconv}
    if acc is not None:
        out["particle_steps"] = int(acc.item())
    if prof_mask:
        ks = {}
        for kid, name in enumerate(_abi.KERNEL_NAMES):
            if not (prof_mask >> kid) & 1:
                continue
            c, t, b = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
            _abi.check(lib.dfsph_profile_read(h, kid, ctypes.byref(c), ctypes.byref(t), ctypes.byref(b)), lib)
            ks[name] = (int(c.value), float(t.value), float(b.value))
        _abi.check(lib.dfsph_set_profiling(h, 0, 1), lib)
        out["kernels"] = ks
    return out


def plain_sample(wl: dict, worlds: list[int], steps: int, warmup: int, dt: float, pair_cache) -> dict:
    """The sampled worlds as plain DFSPH: one single-world solver each (every phase one launch over that world), no
    learned start, run one after another for the same warm-up and step counts."""
    import torch

    from alloy_amd.sph import _abi

    n_tot, ms_tot, den, its = 0, 0.0, [], []
    pour = wl["kind"] == "pour_worlds"
    for w in worlds:
        sc = make_scene(wl, "cuda", subset=[w])
        solver = sc.make_solver()
        if pair_cache is not None:
            solver.set_option(_abi.OPT_PAIR_CACHE, pair_cache)
        state = sc.model.state()
        run = Run(sc, solver, state)
        for _ in range(warmup):
            run.step(dt)
        r = timed_run(run, steps, dt, count_active=pour)
        n_tot += r["particle_steps"] if pour else sc.fluid_counts[0] * steps
        ms_tot += r["total_ms"]
        den.append(r["density_err_mean"])
        its.append((r["iters_density_mean"], r["iters_divergence_mean"]))
        del solver, state, sc, run
        torch.cuda.empty_cache()
    return {"value": n_tot / (ms_tot * 1e-3), "unit": "particle-steps/s", "worlds_sampled": len(worlds),
            "world_ids": worlds, "steps": steps, "warmup": warmup, "ms_per_world_step": ms_tot / (steps * len(worlds)),
            "density_err_mean": float(np.mean([d for d in den if d is not None])),
            "iters_density_mean": float(np.mean([i[0] for i in its])),
            "iters_divergence_mean": float(np.mean([i[1] for i in its])),
            "solver": "plain DFSPH, one single-world solver per world stepped in turn, no learned warm start"}


def cpu_info() -> dict:
    """The host CPU: model name, logical CPUs (nproc) and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(wl: dict, steps: int, warmup: int, sample_w: int) -> dict:
    """The CPU oracle (float32 C++, the same algorithm and operation order) on a bounded sample: the whole scene for
    single-world workloads, an even sample of `sample_w` worlds (each exactly as in the pack) for the batched ones;
    timed on one core (the serial build) and on every core the process may use (the OpenMP build, OMP_NUM_THREADS
    threads, else one per available CPU)."""
    from alloy_amd.sph import _abi, scenes
    from oracle.oracle import OracleDFSPH

    pour = wl["kind"] == "pour_worlds"
    subset = sample_worlds(wl["worlds"], sample_w) if wl["kind"] in ("bowl_worlds", "pour_worlds") else None
    sc = make_scene(wl, "cpu", subset=subset)
    if steps is None:
        steps = wl["steps"] if pour else int(max(2, min(40, CPU_SAMPLE_PARTICLE_STEPS // max(sum(sc.fluid_counts), 1))))
    info = cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", info["affinity"]) or info["affinity"])
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    n = sum(sc.fluid_counts)

    def timed(parallel: bool, k_steps: int) -> tuple[float, float]:
        if pour:  # reset the spares the previous pass emitted
            sc.model.particle_flags.copy_(flags0)
            sc.model.sph.phase.copy_(phase0)
        o = OracleDFSPH(sc.model, sc.config, 32, parallel=parallel)
        step = lambda k: o.step(scenes.DT)  # noqa: E731
        if pour:  # the scene's emitters through the same host bookkeeping, and the swaying walls
            from alloy_amd.sph.emitters import EmitterManager

            def cursors():
                c = np.asarray(o.buffer(_abi.BUF_EMIT_CURSOR)).copy()
                return np.concatenate([c, np.zeros(len(mgr.table.world) - len(c), np.int32)])

            mgr = EmitterManager(sc.model, sc.config, o.set_emitters, cursors)
            for args in sc.emitters:
                mgr.add_emitter(*args)
            drive_state = sc.model.state()
            wall_ids = sc.sway["ids"].numpy()

            def step(k):
                sc.drive(drive_state, k * scenes.DT)
                o.qd[wall_ids] = drive_state.particle_qd.numpy()[wall_ids]
                o.step(scenes.DT)

        for k in range(warmup):
            step(k)
        act = lambda: int(np.count_nonzero(((o.buffer(_abi.BUF_FLAGS) & 1) != 0) & (o.buffer(_abi.BUF_PHASE) >= 1)))  # noqa: E731
        n_ps = 0
        t = time.perf_counter()
        for k in range(warmup, warmup + k_steps):
            step(k)
            n_ps += act() if pour else 0
        dt = time.perf_counter() - t
        return (n_ps if pour else n * k_steps) / dt, dt

    flags0, phase0 = sc.model.particle_flags.clone(), sc.model.sph.phase.clone()
    serial, t1 = timed(False, steps)
    par_steps = steps if pour else steps * max(1, min(threads, 8))
    parallel, tp = timed(True, par_steps)
    what = (f"{len(subset)} of the {wl['worlds']} worlds ({n} {'spares' if pour else 'liquid particles'})" if subset
            else f"{n} particles")
    return {"value": parallel, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (float32 C++ -O3, OpenMP over the per-particle loops, {threads} threads) on {what} "
                      f"of the same workload from its initial state: {warmup} warm-up + {par_steps} timed steps "
                      f"({tp:.1f} s)",
            "serial_value": serial, "serial_cores": 1,
            "serial_sample": f"the serial build on the same sample, {warmup} warm-up + {steps} timed steps ({t1:.1f} s)",
            **info}


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


def options_string(options: dict) -> str:
    return ",".join(f"{k}={v}" for k, v in sorted(options.items()))


# options that leave the roofline window's k_accel / k_apply launches unchanged: the list build's block size, and the
# graph and device-loop paths, which per-launch timing switches off
TRAFFIC_NEUTRAL = ("build_block", "build_staged", "device_loop", "graphs")


def _traffic_options(options: str) -> str:
    return ",".join(kv for kv in options.split(",") if kv.split("=")[0] not in TRAFFIC_NEUTRAL)


def _traffic_record(kernel: str, workload: str, steps: int, warmup: int, options: str):
    """The committed rocprofv3 FETCH_SIZE + WRITE_SIZE summary of `kernel` on `workload`, if it was measured on the
    roofline window of a bench run with these step counts and with options that launch the same `kernel` work;
    (record or None, reason)."""
    name = f"{kernel.split('<')[0]}_{workload}_traffic.json"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, f"no profiles/{name}"
    with open(path) as f:
        recs = json.load(f)
    recs = recs if isinstance(recs, list) else [recs]  # one record per step counts and options
    for rec in recs:
        if (rec.get("steps"), rec.get("warmup")) == (steps, warmup) and \
                _traffic_options(rec.get("options", "")) == _traffic_options(options):
            return rec, f"profiles/{name}: {rec.get('window', '')}; command: {rec.get('command', '')}"
    have = "; ".join(f"--steps {r.get('steps')} --warmup {r.get('warmup')} options {r.get('options')}" for r in recs)
    return None, f"profiles/{name} holds {have}, not this run's --steps {steps} --warmup {warmup} options {options}"


LEG_1M_WARMUP = 60  # past the first steps of the collapse, where every solve sits at min_iterations whatever its start


def learned_1m_leg(net, solves: int, steps: int) -> dict:
    """The 1,000,000-particle dam break (100^3, seed 3, one world, 24-bit rows): `steps` timed steps after
    LEG_1M_WARMUP warm-up steps with the learned start and with the rule, the same tolerances; ms per step and
    iterations."""
    import torch

    from alloy_amd.sph import _abi, scenes

    out = {"workload": "dam break, 1,000,000 particles (100^3 lattice, seed 3) in a 2 m SDF box tank, 1 world",
           "steps": steps, "warmup": LEG_1M_WARMUP}
    for name, use in (("rule", None), ("learned", net)):
        _log(f"[bench] 1M dam break, {name} start")
        sc = scenes.dam_break(100, 2.0, 3, device="cuda")
        s = sc.make_solver()
        if use is not None:
            use.attach(s)
            s.set_option(_abi.OPT_LEARNED_SOLVES, solves)
        st = sc.model.state()
        run = Run(sc, s, st)
        for _ in range(LEG_1M_WARMUP):
            run.step(scenes.DT)
        r = timed_run(run, steps, scenes.DT)
        out[name] = {"ms_per_step": r["total_ms"] / steps, "value": 1.0e6 * steps / (r["total_ms"] * 1e-3),
                     "iters_density_mean": r["iters_density_mean"], "iters_divergence_mean": r["iters_divergence_mean"],
                     "density_err_mean": r["density_err_mean"]}
        del s, st, run, sc
        torch.cuda.empty_cache()
    out["learned_vs_rule"] = out["learned"]["value"] / out["rule"]["value"]
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="bowl_worlds_512", choices=sorted(WORKLOADS))
    ap.add_argument("--learned", default="auto", choices=["auto", "on", "off"],
                    help="learned warm start (auto: on when the trained weights are present)")
    ap.add_argument("--plain-worlds", type=int, default=16, help="worlds sampled for the plain single-world run")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-rule-compare", action="store_true",
                    help="skip the batched run with the rule warm start (the learned start's own effect)")
    ap.add_argument("--check-every", type=int, default=1, help="iterations between host convergence checks")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the autotuner (it runs after the warm-up, outside the timed region, and changes no "
                         "result bit)")
    ap.add_argument("--graphs", action="store_true", help="replay the step's segments as hipGraphs")
    ap.add_argument("--pair-cache", type=int, default=None, help="DFSPH_PAIR_CACHE_* mask (default: the library's)")
    ap.add_argument("--options", default="",
                    help="name=value,... of library options set before the warm-up (e.g. the autotuner's choice, to "
                         "profile exactly the timed configuration with --no-autotune)")
    ap.add_argument("--no-single-launch", action="store_true",
                    help="skip the plain single-launch run (rule start, default options) behind vs_single_launch")
    ap.add_argument("--learned-1m-steps", type=int, default=40,
                    help="steps of the 1,000,000-particle dam-break leg timing the learned start against the rule "
                         "(0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="timed oracle steps (default: about 3e6 particle-steps of the sample)")
    ap.add_argument("--cpu-worlds", type=int, default=4, help="worlds of the batched pack in the CPU sample")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--roofline-sample", type=int, default=1,
                    help="time one in this many k_accel / k_apply launches of the roofline window (pseudo-random)")
    ap.add_argument("--breakdown-steps", type=int, default=30,
                    help="steps after the timed region with every kernel timed (per-kernel table)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB (3 launches) and exit: the known-size "
                         "kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this benchmark measures one GPU; --gpus {args.gpus} is not supported (the batched packs "
              "shard whole worlds across GPUs with no exchange, so a pack's rate scales with its GPU count)",
              file=sys.stderr)
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
    if args.steps is None:
        args.steps = wl["steps"]
    dt = scenes.DT
    _log(f"[bench] workload {args.workload} {wl} steps={args.steps} warmup={args.warmup}")
    sc = make_scene(wl, "cuda")
    flags0, phase0 = sc.model.particle_flags.clone(), sc.model.sph.phase.clone()
    solver = sc.make_solver()
    solver.set_check_interval(args.check_every)
    solver.set_graphs(args.graphs)
    if args.pair_cache is not None:
        solver.set_option(_abi.OPT_PAIR_CACHE, args.pair_cache)
    net, solves = load_learned(args.learned, "cuda")
    if net is not None:
        net.attach(solver)
        solver.set_option(_abi.OPT_LEARNED_SOLVES, solves)
    for item in filter(None, args.options.split(",")):
        name, v = item.split("=")
        solver.set_option(OPTION_IDS[name.strip()], int(v))
    state = sc.model.state()
    pour = wl["kind"] == "pour_worlds"
    n_fluid = sum(sc.fluid_counts)
    W = sc.model.world_count
    main_run = Run(sc, solver, state)
    for _ in range(args.warmup):
        main_run.step(dt)
    torch.cuda.synchronize()
    _log("[bench] warm-up done")
    tuned = None
    if not args.no_autotune:
        from alloy_amd.sph.autotune import autotune

        tuned = autotune(solver, state, dt, learned=True, log=_log)
    learned_used = net is not None and solver.get_option(_abi.OPT_LEARNED_SOLVES) != 0
    options_used = {name: solver.get_option(oid) for name, oid in OPTION_IDS.items()}
    # the timed region: profiling off, exactly the configuration the options above describe
    run = timed_run(main_run, args.steps, dt, count_active=pour)
    total_ms = run["total_ms"]
    _log(f"[bench] timed {args.steps} steps in {total_ms:.1f} ms")
    particle_steps = run["particle_steps"] if pour else n_fluid * args.steps
    value = particle_steps / (total_ms * 1e-3)
    # the roofline window: the next `steps` steps, every k_accel / k_apply launch timed inside its own dispatch and its
    # algorithmic bytes counted by the kernels (per-launch events cost the host, so this is not the timed region)
    roof_mask = (1 << _abi.KID_ACCEL) | (1 << _abi.KID_APPLY)
    roof = timed_run(main_run, args.steps, dt, prof_mask=roof_mask, sample=args.roofline_sample, count_active=pour)
    ks = roof["kernels"]
    dom = max(ks, key=lambda k: ks[k][1])
    nl, tl, bl = ks[dom]
    achieved = bl / (tl * 1e-3) / 1e9
    w0 = args.warmup + args.steps
    traffic_rec = _traffic_record(dom, args.workload, args.steps, args.warmup, options_string(options_used))
    solver_desc = ("batched, one launch per phase over all worlds" if W > 1 else "single world") + (
        f", learned warm start (trained network; solves {'divergence+density' if solves == 3 else 'density' if solves == 2 else 'divergence'})"
        if learned_used else ", the reference's warm-start rule" + (
            " (the autotuner timed the learned start slower on this pack)" if net is not None else ""))
    if options_used.get("device_loop"):
        solver_desc += ", device-side solve loop (no host round trip in a step)"
    out = {
        "metric": METRIC,
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
        "config": {"workload": describe(wl, sc), "particles": n_fluid, "worlds": W, "solver": solver_desc},
        "particle_steps": particle_steps,
        "regime": {"window": f"steps {args.warmup + 1}-{w0} of the run (warm-up {args.warmup} steps, autotuning "
                             "restored to the warm-up's end state)",
                   "iters_density_mean": run["iters_density_mean"],
                   "iters_divergence_mean": run["iters_divergence_mean"],
                   "density_err_mean": run["density_err_mean"],
                   "density_err_max_sampled": run["density_err_max_sampled"],
                   "unconverged_sampled_steps": run["unconverged_sampled_steps"],
                   "note": "iterations: per-world means over every step of the window; density error: per-world "
                           "means over the reports every 10 steps; a world's solve stops at its tolerance or at "
                           "min_iterations = 2"},
        "density_err_mean": run["density_err_mean"],
        "iters_density_mean": run["iters_density_mean"],
        "iters_divergence_mean": run["iters_divergence_mean"],
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_rec[0]["hbm_bytes_per_launch"] if traffic_rec[0] else None,
                     "alg_bytes_per_launch": bl / max(nl, 1), "avg_launch_ms": tl / max(nl, 1), "launches_timed": nl,
                     "window": f"steps {w0 + 1}-{w0 + args.steps} (the {args.steps} steps after the timed region), "
                               f"1 in {args.roofline_sample} launches of {dom} timed by HIP events inside its "
                               "dispatch, algorithmic bytes counted by the kernel" + (
                                   "; per-launch timing runs the solve's phases as separate launches, so the "
                                   "device-side loop is off in this window" if options_used.get("device_loop") else ""),
                     "traffic_source": traffic_rec[1]},
        "kernel_ms_per_step_roofline_window": {k: round(v[1] / args.steps, 4) for k, v in ks.items()},
        "options": options_used,
        "tuning": {"autotuned": tuned is not None,
                   "ms_per_step": tuned["ms_per_step"] if tuned else None,
                   "seconds": round(tuned["seconds"], 2) if tuned else 0.0},
    }
    if args.breakdown_steps > 0:
        bd = timed_run(main_run, args.breakdown_steps, dt, prof_mask=(1 << len(_abi.KERNEL_NAMES)) - 1)
        out["kernel_ms_per_step"] = {k: round(v[1] / args.breakdown_steps, 4) for k, v in bd["kernels"].items()}
        out["kernel_launches_per_step"] = {k: round(v[0] / args.breakdown_steps, 2) for k, v in bd["kernels"].items()}
        out["breakdown_steps"] = args.breakdown_steps
    if not args.no_fp64:
        torch.cuda.synchronize()
        _log("[bench] FP64 density error of the final state")
        out["fp64_density_err"] = fp64_density_error(sc, state.particle_q.cpu().numpy(),
                                                     state.particle_qd.cpu().numpy())
        _log(f"[bench] fp64 density error {out['fp64_density_err']:.3e}")
    del solver, state, main_run
    torch.cuda.empty_cache()

    def rerun(options: dict, use_net) -> dict:
        """The same pack from its initial state with these options and start, the same warm-up and timed steps."""
        if pour:  # the first run emitted from the model's spares: back to the build-time flags and phase
            sc.model.particle_flags.copy_(flags0)
            sc.model.sph.phase.copy_(phase0)
        s2 = sc.make_solver()
        for name, v in options.items():
            if name != "learned_solves":
                try:
                    s2.set_option(OPTION_IDS[name], v)
                except ValueError:
                    pass
        if use_net is not None:
            use_net.attach(s2)
            s2.set_option(_abi.OPT_LEARNED_SOLVES, solves)
        st2 = sc.model.state()
        run2 = Run(sc, s2, st2)
        for _ in range(args.warmup):
            run2.step(dt)
        r2 = timed_run(run2, args.steps, dt, count_active=pour)
        ps2 = r2["particle_steps"] if pour else n_fluid * args.steps
        res = {"value": ps2 / (r2["total_ms"] * 1e-3), "ms_per_step": r2["total_ms"] / args.steps,
               "iters_density_mean": r2["iters_density_mean"], "iters_divergence_mean": r2["iters_divergence_mean"],
               "density_err_mean": r2["density_err_mean"]}
        del s2, st2, run2
        torch.cuda.empty_cache()
        return res

    if net is not None and not args.no_rule_compare:
        # the other warm start on the same pack with the same options: the learned start's own effect
        if learned_used:
            _log("[bench] the same run with the rule warm start")
            other = rerun(options_used, None)
            out["batched_rule_start"] = other
            out["learned_vs_rule_start"] = value / other["value"]
        else:
            _log("[bench] the same run with the learned warm start")
            other = rerun(options_used, net)
            out["batched_learned_start"] = other
            out["learned_vs_rule_start"] = other["value"] / value
        out["learned_vs_rule_note"] = ("learned / rule warm start, batched, same options and steps; the headline uses "
                                       "the start the autotuner timed faster")
    if not args.no_single_launch:
        # the plain single-launch solver: all worlds in one launch per phase (the reference's own structure), the
        # reference's rule start, the library's default options, no autotuning
        _log("[bench] plain single-launch DFSPH (rule start, default options)")
        single = rerun({}, None)
        out["single_launch_rule_start"] = single
        out["vs_single_launch"] = value / single["value"]
        _log(f"[bench] vs single launch {out['vs_single_launch']:.3f}")
    if args.learned_1m_steps > 0 and net is not None:
        out["dam_break_1m_learned_vs_rule"] = learned_1m_leg(net, solves, args.learned_1m_steps)
    if wl["kind"] in ("bowl_worlds", "pour_worlds") and not args.no_plain:
        _log(f"[bench] plain single-world DFSPH on {args.plain_worlds} sampled worlds")
        plain = plain_sample(wl, sample_worlds(W, args.plain_worlds), args.steps, args.warmup, dt, args.pair_cache)
        plain["note"] = ("one single-world solver per sampled world, stepped one after another: the per-launch cost "
                         "of small worlds, which the reference avoids by launching all worlds at once (that structure "
                         "is single_launch_rule_start)")
        out["plain"] = plain
        out["speedup_vs_plain"] = value / plain["value"]
        _log(f"[bench] plain {plain['value']:.3e} particle-steps/s, speedup {out['speedup_vs_plain']:.2f}")
    if not args.no_cpu_baseline:
        _log("[bench] timing the CPU oracle")
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_steps, 2, args.cpu_worlds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
