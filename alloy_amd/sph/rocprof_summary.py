"""Summaries of rocprofv3 output for the measurement records under profiles/ (run where the profile was taken).

    python -m alloy_amd.sph.rocprof_summary kernels DIR --command CMD --title T --out FILE.md [--window A-B]
        kernel trace (rocprofv3 --kernel-trace -f csv): calls, mean and total duration and share per kernel; with
        --window A-B[,C-D...], the same for the dispatches of steps A..B of the run alone (a step starts at its
        k_groups_hist)
    python -m alloy_amd.sph.rocprof_summary traffic FETCH_DIR WRITE_DIR --kernel K --window A-B --fetch-scale F
            --write-scale S --steps K --warmup W --options O --command CMD --out FILE.json
        HBM bytes per launch of kernel K over its dispatches in steps A..B (the bench's roofline window) from two
        --pmc runs (FETCH_SIZE, WRITE_SIZE in KiB per dispatch, k_groups_hist included to mark the steps), each
        scaled by what the counter reads for a plain copy kernel of known size; with --append the file keeps a list
        of such records, one per step counts and options
    python -m alloy_amd.sph.rocprof_summary counters DIR --kernels A,B --top FRAC --command CMD --out FILE.md
        mean per dispatch of every --pmc counter of the named kernels, over their dispatches doing at least FRAC of the
        largest one's work (the launches in which every world still iterates)

Kernel names are shortened to their template form (k_apply<false>) without the namespace and argument list.
"""

from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os


def _short(name: str) -> str:
    n = name.replace("void ", "").replace("(anonymous namespace)::", "")
    return n.split("(")[0]


def _rows(d: str, suffix: str) -> list[dict]:
    rows = []
    for f in sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    if not rows:
        raise SystemExit(f"no *{suffix} under {d}")
    return rows


STEP_KERNEL = "k_groups_hist"  # the first kernel of every step's neighbour phase


def _window(order: list[tuple[int, str]], window: str) -> set[int]:
    """Dispatch ids of steps A..B (1-based) of the run, from the dispatch order (id, short name)."""
    a, b = (int(x) for x in window.split("-"))
    starts = [i for i, (_, k) in enumerate(order) if k.startswith(STEP_KERNEL)]
    if len(starts) < b:
        raise SystemExit(f"the run has {len(starts)} steps, window {window}")
    lo = starts[a - 1]
    hi = starts[b] if b < len(starts) else len(order)
    return {order[i][0] for i in range(lo, hi)}


def _table(agg, title_lines):
    total = sum(v[1] for v in agg.values())
    lines = list(title_lines) + ["| kernel | calls | average us | total ms | share |", "|---|---|---|---|---|"]
    for k, (n, ns) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:40]:
        lines.append(f"| {k} | {n} | {ns / n / 1e3:.2f} | {ns / 1e6:.2f} | {100.0 * ns / total:.2f}% |")
    lines += ["", f"All kernels: {sum(v[0] for v in agg.values())} dispatches, {total / 1e6:.2f} ms.", ""]
    return lines


def kernels(args) -> None:
    rows = _rows(args.dir, "kernel_trace.csv")
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    agg = collections.defaultdict(lambda: [0, 0])
    for r in rows:
        k = _short(r["Kernel_Name"])
        agg[k][0] += 1
        agg[k][1] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    lines = [f"# {args.title}", "", f"Command: `{args.command}`", ""]
    if args.note:
        lines += [args.note, ""]
    lines += _table(agg, ["## The whole run", ""])
    order = [(int(r["Dispatch_Id"]), _short(r["Kernel_Name"])) for r in rows]
    for window in filter(None, args.window.split(",")):
        ids = _window(order, window)
        win = collections.defaultdict(lambda: [0, 0])
        for r in rows:
            if int(r["Dispatch_Id"]) in ids:
                k = _short(r["Kernel_Name"])
                win[k][0] += 1
                win[k][1] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        lines += _table(win, [f"## Steps {window} alone", ""])
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")


def _per_dispatch(rows: list[dict], kernel: str) -> dict[int, dict[str, float]]:
    per = collections.defaultdict(dict)
    for r in rows:
        if not _short(r["Kernel_Name"]).startswith(kernel.split("<")[0]):
            continue
        per[int(r["Dispatch_Id"])][r["Counter_Name"]] = float(r["Counter_Value"])
    return per


def _window_ids(rows: list[dict], window: str) -> set[int]:
    order = sorted({(int(r["Dispatch_Id"]), _short(r["Kernel_Name"])) for r in rows})
    return _window(order, window)


def traffic(args) -> None:
    frows = _rows(args.fetch_dir, "counter_collection.csv")
    wrows = _rows(args.write_dir, "counter_collection.csv")
    fwin, wwin = _window_ids(frows, args.window), _window_ids(wrows, args.window)
    fetch = _per_dispatch(frows, args.kernel)
    write = _per_dispatch(wrows, args.kernel)
    fid, wid = sorted(i for i in fetch if i in fwin), sorted(i for i in write if i in wwin)
    f_kib = sum(fetch[i]["FETCH_SIZE"] for i in fid) / len(fid)
    w_kib = sum(write[i]["WRITE_SIZE"] for i in wid) / len(wid)
    rec = {"kernel": args.kernel, "workload": args.workload, "steps": args.steps, "warmup": args.warmup,
           "options": args.options, "command": args.command,
           "window": f"the {len(fid)} dispatches of {args.kernel} in steps {args.window} of the run: the bench's "
                     "roofline window",
           "fetch_kib_per_launch": f_kib, "write_kib_per_launch": w_kib,
           "fetch_scale": args.fetch_scale, "write_scale": args.write_scale,
           "hbm_bytes_per_launch": 1024.0 * (f_kib * args.fetch_scale + w_kib * args.write_scale)}
    out = rec
    if args.append and os.path.exists(args.out):  # one record per step counts and options: replace or add
        with open(args.out) as f:
            old = json.load(f)
        recs = old if isinstance(old, list) else [old]
        same = lambda r: (r.get("steps"), r.get("warmup"), r.get("options")) == (rec["steps"], rec["warmup"],
                                                                                   rec["options"])
        out = [r for r in recs if not same(r)] + [rec]
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


def _split_names(spec: str) -> list[str]:
    """Kernel names separated by commas outside template brackets ("k_a<false>,k_b<false, 128, false>")."""
    out, cur, depth = [], "", 0
    for ch in spec:
        depth += (ch == "<") - (ch == ">")
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def counters(args) -> None:
    rows = _rows(args.dir, "counter_collection.csv")
    lines = [f"# {args.title}", "", f"Command: `{args.command}`", ""]
    if args.note:
        lines += [args.note, ""]
    for k in _split_names(args.kernels):
        per = _per_dispatch(rows, k)
        if not per:
            continue
        names = sorted({c for d in per.values() for c in d})
        key = "SQ_INSTS_VALU" if "SQ_INSTS_VALU" in names else names[0]
        top = max(d.get(key, 0.0) for d in per.values())
        ids = [i for i, d in per.items() if d.get(key, 0.0) >= args.top * top]
        lines += [f"## {k}", "", f"{len(ids)} of {len(per)} dispatches (those with at least {args.top:g} of the "
                  f"largest {key}); mean per dispatch:", "", "| counter | mean |", "|---|---|"]
        for c in names:
            lines.append(f"| {c} | {sum(per[i].get(c, 0.0) for i in ids) / len(ids):.6g} |")
        waves = sum(per[i].get("SQ_WAVES", 0.0) for i in ids) / len(ids)
        if waves:
            lines += ["", "Per wave: " + ", ".join(
                f"{c} {sum(per[i].get(c, 0.0) for i in ids) / len(ids) / waves:.1f}"
                for c in names if c.startswith("SQ_INSTS"))]
        lines.append("")
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")


def main() -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("kernels")
    a.add_argument("dir")
    a.add_argument("--title", required=True)
    a.add_argument("--command", required=True)
    a.add_argument("--note", default="")
    a.add_argument("--window", default="")
    a.add_argument("--out", required=True)
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--kernel", required=True)
    b.add_argument("--workload", required=True)
    b.add_argument("--window", required=True)
    b.add_argument("--fetch-scale", type=float, required=True)
    b.add_argument("--write-scale", type=float, required=True)
    b.add_argument("--steps", type=int, required=True)
    b.add_argument("--warmup", type=int, required=True)
    b.add_argument("--options", default="")
    b.add_argument("--command", required=True)
    b.add_argument("--out", required=True)
    b.add_argument("--append", action="store_true", help="keep the file's records of other step counts or options")
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--kernels", required=True)
    c.add_argument("--top", type=float, default=0.8)
    c.add_argument("--title", required=True)
    c.add_argument("--command", required=True)
    c.add_argument("--note", default="")
    c.add_argument("--out", required=True)
    args = ap.parse_args()
    {"kernels": kernels, "traffic": traffic, "counters": counters}[args.cmd](args)


if __name__ == "__main__":
    main()
