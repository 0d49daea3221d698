"""Summaries of rocprofv3 CSV output (``-f csv``) for the kernels of this
library: per-kernel durations from a kernel trace, and HBM bytes per
dispatch from FETCH_SIZE / WRITE_SIZE counter runs scaled by what the same
counters read for a plain device copy of known size.

    python -m ereb.rocprof_summary kernels TRACE_DIR --title T --command C --out OUT.md
    python -m ereb.rocprof_summary traffic --fetch D --write D --copy-fetch D \\
        --copy-write D --copy-bytes N --workload W --form F --command C \\
        --out-json OUT.json --out-md OUT.md
    python -m ereb.rocprof_summary counters D [D ...] --kernel PREFIX [...] \\
        --title T --command C [--command C ...] --out OUT.md
    python -m ereb.rocprof_summary timeline TRACE_DIR --first PREFIX --index N \\
        --title T --command C --out OUT.md

TRACE_DIR and the counter directories are rocprofv3 output directories
(searched recursively for ``*kernel_trace.csv`` / ``*counter_collection.csv``).
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
from collections import defaultdict

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def _rows(d: str, suffix: str):
    files = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    for f in files:
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                yield {k.strip().lower(): v for k, v in r.items() if k}


def _grid(r) -> str:
    gx = r.get("grid_size_x") or r.get("grid_size") or "?"
    gy = r.get("grid_size_y", "1")
    wx = r.get("workgroup_size_x") or r.get("workgroup_size") or "?"
    try:
        return f"({int(gx) // max(int(wx), 1)}, {int(gy)}) x {wx}"
    except ValueError:
        return f"{gx} / {wx}"


def kernel_table(trace_dir: str):
    """[(name, grid, calls, mean_us, total_us)] sorted by total time."""
    acc = defaultdict(list)
    for r in _rows(trace_dir, "kernel_trace.csv"):
        t = (int(r["end_timestamp"]) - int(r["start_timestamp"])) / 1e3
        acc[(r["kernel_name"], _grid(r))].append(t)
    out = [(k[0], k[1], len(v), sum(v) / len(v), sum(v)) for k, v in acc.items()]
    return sorted(out, key=lambda x: -x[4])


def _short(name: str, n: int = 90) -> str:
    return name if len(name) <= n else name[:n - 3] + "..."


def write_kernels(trace_dir, title, command, out, top=40):
    rows = kernel_table(trace_dir)
    total = sum(r[4] for r in rows)
    lines = [f"# {title}", "", "Command (run from /tmp with TMPDIR=/tmp; `$R` = repository root):",
             "", f"    {command}", "",
             "Durations from the kernel trace (end - start per dispatch), grouped by kernel "
             "and launch grid (workgroups x workgroup size).", "",
             "| kernel | grid x block | calls | mean µs | total µs | % |",
             "|---|---|---:|---:|---:|---:|"]
    for name, grid, calls, mean, tot in rows[:top]:
        lines.append(f"| `{_short(name)}` | {grid} | {calls} | {mean:.2f} | {tot:.1f} | "
                     f"{100 * tot / total:.2f} |")
    open(out, "w").write("\n".join(lines) + "\n")


def counter_means(d: str, counter: str):
    """{kernel name: mean counter value per dispatch} (summed over the
    counter's instances within a dispatch)."""
    per = defaultdict(float)
    names = {}
    for r in _rows(d, "counter_collection.csv"):
        if r.get("counter_name", "").strip() != counter:
            continue
        key = (r.get("dispatch_id"), r.get("correlation_id"))
        per[key] += float(r["counter_value"])
        names[key] = r["kernel_name"]
    acc = defaultdict(list)
    for key, v in per.items():
        acc[names[key]].append(v)
    return {k: sum(v) / len(v) for k, v in acc.items()}


def write_traffic(a):
    fetch = counter_means(a.fetch, "FETCH_SIZE")
    write = counter_means(a.write, "WRITE_SIZE")
    cf = counter_means(a.copy_fetch, "FETCH_SIZE")
    cw = counter_means(a.copy_write, "WRITE_SIZE")
    ck = [k for k in cf if k.startswith(COPY_KERNEL)]
    if not ck:
        raise RuntimeError("no copy kernel in the calibration runs")
    # the largest dispatch of the copy kernel is the calibration copy
    fk = max(ck, key=lambda k: cf[k])
    fscale = a.copy_bytes / (cf[fk] * 1024.0)
    wscale = a.copy_bytes / (max(cw[k] for k in cw if k.startswith(COPY_KERNEL)) * 1024.0)
    kern = {}
    for k in sorted(set(fetch) | set(write)):
        if not k.startswith("void ereb::") and not k.startswith("ereb::"):
            continue
        fkib, wkib = fetch.get(k, 0.0), write.get(k, 0.0)
        kern[k] = {"fetch_kib": fkib, "write_kib": wkib,
                   "bytes_per_launch": fkib * 1024 * fscale + wkib * 1024 * wscale}
    res = {"workload": a.workload, "form": a.form, "copy_kernel": COPY_KERNEL,
           "copy_bytes": a.copy_bytes, "fetch_scale": fscale, "write_scale": wscale,
           "command": a.command, "kernels": kern}
    json.dump(res, open(a.out_json, "w"), indent=1)
    lines = [f"# HBM traffic per dispatch, {a.workload} workload, {a.form} form", "",
             "Commands (each its own run, from /tmp with TMPDIR=/tmp; `$R` = repository root):",
             "", f"    {a.command}", "",
             f"Calibration: a plain copy of {a.copy_bytes} bytes (`{COPY_KERNEL}`) reads "
             f"FETCH_SIZE x {fscale:.4f} and WRITE_SIZE x {wscale:.4f} bytes per KiB counted, so "
             "HBM bytes = FETCH_SIZE x 1024 x fetch_scale + WRITE_SIZE x 1024 x write_scale.", "",
             "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes (scaled) |",
             "|---|---:|---:|---:|"]
    for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["bytes_per_launch"]):
        lines.append(f"| `{_short(k)}` | {v['fetch_kib']:.0f} | {v['write_kib']:.0f} | "
                     f"{v['bytes_per_launch']:.4g} |")
    open(a.out_md, "w").write("\n".join(lines) + "\n")


DERIVED = [
    ("waves", lambda c: c.get("SQ_WAVES")),
    ("VALU instr per wave", lambda c: _div(c.get("SQ_INSTS_VALU"), c.get("SQ_WAVES"))),
    ("VALU active", lambda c: _div(c.get("SQ_ACTIVE_INST_VALU"), c.get("SQ_WAVE_CYCLES"))),
    ("waiting on any", lambda c: _div(c.get("SQ_WAIT_INST_ANY"), c.get("SQ_WAVE_CYCLES"))),
    ("LDS instr per wave", lambda c: _div(c.get("SQ_INSTS_LDS"), c.get("SQ_WAVES"))),
    ("LDS bank-conflict share",
     lambda c: _div(c.get("SQ_LDS_BANK_CONFLICT"), c.get("SQ_ACTIVE_INST_LDS"))),
    ("L2 hit rate", lambda c: _div(c.get("TCC_HIT"), _add(c.get("TCC_HIT"), c.get("TCC_MISS")))),
]


def _div(a, b):
    return None if a is None or not b else a / b


def _add(a, b):
    return None if a is None or b is None else a + b


def write_counters(dirs, kernels, title, commands, out):
    """Per-dispatch means of every counter found in the runs `dirs`, for the
    kernels whose names start with one of `kernels`, and derived rates."""
    vals = defaultdict(dict)
    for d in dirs:
        rows = list(_rows(d, "counter_collection.csv"))
        for cname in sorted({r["counter_name"].strip() for r in rows}):
            base = cname.replace("_sum", "")
            for k, v in counter_means(d, cname).items():
                vals[k][base] = v
    lines = [f"# {title}", "", "Commands (each its own run, from /tmp with TMPDIR=/tmp; "
             "`$R` = repository root):", ""] + [f"    {c}" for c in commands] + [
             "", "Per-dispatch means.  Cycle counters are summed over waves, so the fractions "
             "are per wave (a wave shares its SIMD with the others resident there); the LDS "
             "bank-conflict share is `SQ_LDS_BANK_CONFLICT / SQ_ACTIVE_INST_LDS`.", "",
             "| kernel | " + " | ".join(n for n, _ in DERIVED) + " |",
             "|---|" + "---:|" * len(DERIVED)]
    for pre in kernels:
        for k in sorted(vals):
            if k.startswith(pre):
                cells = []
                for _, f in DERIVED:
                    v = f(vals[k])
                    cells.append("—" if v is None else (f"{v:.0f}" if v >= 100 else f"{v:.3g}"))
                lines.append(f"| `{_short(k, 60)}` | " + " | ".join(cells) + " |")
    open(out, "w").write("\n".join(lines) + "\n")


def step_timeline(trace_dir: str, first: str, index: int):
    """Dispatches from the index-th dispatch whose name starts with `first`
    up to the next one, in start order: [(start_us, end_us, queue, name)]
    relative to the first dispatch's start."""
    rows = sorted(((int(r["start_timestamp"]), int(r["end_timestamp"]),
                    r.get("queue_id", "?"), r["kernel_name"])
                   for r in _rows(trace_dir, "kernel_trace.csv")), key=lambda x: x[0])
    marks = [i for i, r in enumerate(rows) if r[3].startswith(first)]
    if index + 1 >= len(marks):
        raise ValueError(f"fewer than {index + 2} dispatches of {first!r}")
    seg = rows[marks[index]:marks[index + 1]]
    t0 = seg[0][0]
    return [((a - t0) / 1e3, (b - t0) / 1e3, q, n) for a, b, q, n in seg]


def write_timeline(trace_dir, first, index, title, command, out):
    tl = step_timeline(trace_dir, first, index)
    span = max(b for _, b, _, _ in tl)
    busy = sum(b - a for a, b, _, _ in tl)
    lines = [f"# {title}", "", "Command (run from /tmp with TMPDIR=/tmp; `$R` = repository root):",
             "", f"    {command}", "",
             f"One step: dispatch {index} of `{_short(first, 60)}` up to the next one, "
             "times in µs from its start, by hardware queue.  Span "
             f"{span:.1f} µs, sum of the durations {busy:.1f} µs.", "",
             "| start | end | duration | queue | kernel |", "|---:|---:|---:|---:|---|"]
    for a, b, q, n in tl:
        lines.append(f"| {a:.1f} | {b:.1f} | {b - a:.1f} | {q} | `{_short(n, 70)}` |")
    open(out, "w").write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("trace_dir")
    k.add_argument("--title", required=True)
    k.add_argument("--command", required=True)
    k.add_argument("--out", required=True)
    t = sub.add_parser("traffic")
    for n in ("--fetch", "--write", "--copy-fetch", "--copy-write", "--workload", "--form",
              "--command", "--out-json", "--out-md"):
        t.add_argument(n, required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--kernel", action="append", required=True, help="kernel name prefix")
    c.add_argument("--title", required=True)
    c.add_argument("--command", action="append", required=True)
    c.add_argument("--out", required=True)
    tl = sub.add_parser("timeline")
    tl.add_argument("trace_dir")
    tl.add_argument("--first", required=True, help="name prefix of a step's first kernel")
    tl.add_argument("--index", type=int, required=True)
    tl.add_argument("--title", required=True)
    tl.add_argument("--command", required=True)
    tl.add_argument("--out", required=True)
    a = ap.parse_args()
    if a.cmd == "timeline":
        write_timeline(a.trace_dir, a.first, a.index, a.title, a.command, a.out)
    elif a.cmd == "kernels":
        write_kernels(a.trace_dir, a.title, a.command, a.out)
    elif a.cmd == "counters":
        write_counters(a.dirs, a.kernel, a.title, a.command, a.out)
    else:
        write_traffic(a)


if __name__ == "__main__":
    main()
