"""Summaries of rocprofv3 CSV output for the solver's kernels.

    python -m vlp2d_amd.rocprof_summary trace <kernel_trace.csv> [<kernel_stats.csv>]
    python -m vlp2d_amd.rocprof_summary pmc <fetch counter_collection.csv> \
        <write counter_collection.csv> --copy-bytes B
    python -m vlp2d_amd.rocprof_summary solve <fetch counter_collection.csv> \
        <write counter_collection.csv> --copy-bytes B --label LABEL

``trace`` groups kernel dispatches by kernel and grid size (launches on
different multigrid levels differ in grid size) and prints launches, average
and total duration as a Markdown table.

``pmc`` averages FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel and
grid size, calibrates both against the runtime's device-to-device copy
kernel of known size (``--copy-bytes``, the bytes the copy reads and also
writes), and prints the calibrated HBM traffic per launch.  For the launch
of each kernel class that moves the most bytes (level 0, whose grid holds four
times the nodes of the next level) it adds a line

    TRAFFIC <class>@level0/<f32|f64> <bytes>

where <class> is the launch class the library reports in its kernel timings.

``solve`` sums the calibrated traffic of every dispatch of the last solve in
the run (``bench.py --only <precision> --steps 1 --warmup 1``: the dispatches
after the last device copy, which restores the iterate before each solve, up
to the end) and prints the per-kernel breakdown and a line

    SOLVE_TRAFFIC <label> <bytes> <dispatches>

with the HBM bytes one whole solve moves.
"""
import argparse
import collections
import csv
import re
import sys

# kernel function name -> launch class reported by vlp2d_mg_kernel_times
KERNEL_CLASS = {
    "k_sdown": "down_fused",
    "k_sdown_r64": "down_fused_resnorm64",
    "k_sup": "up_fused",
    "k_lt_up": "up_lds",
    "k_coarse": "coarse_vcycle",
    "k_refine_stream": "update_mixed",
    "k_refine": "update_mixed",
    "k_update_mixed": "update_mixed",
    "k_update_oop": "update_mixed",
    "k_resnorm64": "resnorm64",
    "k_reduce": "reduce",
    "k_finish": "reduce",
    "k_rb_half": "rb_half",
    "k_residual": "residual",
    "k_restrict": "restrict",
    "k_prolong": "prolong",
    "k_fill": "fill",
    "k_copy_interior": "copy_interior",
    "k_absdiff_max": "absdiff_max",
    "k_tdown": "adjoint_down_fused",
    "k_tup": "adjoint_up_fused",
    "k_adj_gather": "adjoint",
    "k_adj_color": "adjoint",
    "k_adj_prolongT": "adjoint",
    "k_adj_restrictT": "adjoint",
    "k_adj_residualT": "adjoint",
    "k_adj_output": "adjoint_output",
}

# classes whose largest grid is the level-0 (finest grid) launch
LEVEL0_CLASSES = {"down_fused", "down_fused_resnorm64", "up_fused", "up_fused_refine", "up_down_fused", "update_mixed",
                  "resnorm64", "rb_half", "residual", "restrict", "prolong", "copy_interior", "absdiff_max"}

COPY_PATTERN = re.compile(r"copyBuffer", re.I)


def short_name(full):
    """'void vlp2d::k_down<float, 2, true>(vlp2d::LevelView<float>, ...)' ->
    'k_down<float, 2, true>'."""
    name = full.split("(", 1)[0]
    name = re.sub(r"^void\s+", "", name).strip()
    name = name.replace("vlp2d::", "")
    return name


def base_and_precision(short):
    base = short.split("<", 1)[0]
    args = short[len(base):]
    prec = "f32" if "float" in args.split(",", 1)[0] or base == "k_sdown_r64" else "f64"
    return base, prec


def kernel_class(short):
    """Launch class of a kernel: the ascent with the mixed-precision
    refinement epilogue (k_sup<T, NS, 3, ...>) and the ascent continued into
    the next cycle's descent (k_sup<T, NS, 4, ...>) are classes of their own;
    k_refine_stream<false> is the residual norm, <true> the update."""
    base, _ = base_and_precision(short)
    targs = [a.strip() for a in short[len(base) + 1:].rstrip(">").split(",")]
    if base == "k_sup" and len(targs) >= 3:
        if targs[2] == "3":
            return "up_fused_refine"
        if targs[2] == "4":
            return "up_down_fused"
    if base == "k_refine_stream" and targs and targs[0] == "false":
        return "resnorm64"
    return KERNEL_CLASS.get(base)


def grid_of(row):
    if "Grid_Size" in row and row["Grid_Size"]:
        return int(row["Grid_Size"])
    return int(row.get("Grid_Size_X", 1)) * int(row.get("Grid_Size_Y", 1)) * int(row.get("Grid_Size_Z", 1))


def trace_table(path, top=40):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        key = (short_name(r["Kernel_Name"]), grid_of(r))
        d[key].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0)
    tot = sum(sum(v) for v in d.values())
    out = ["| kernel | grid (threads) | launches | avg us | total us | share |",
           "|---|---|---|---|---|---|"]
    for (name, grid), v in sorted(d.items(), key=lambda kv: -sum(kv[1]))[:top]:
        out.append(f"| `{name}` | {grid} | {len(v)} | {sum(v) / len(v):.1f} | {sum(v):.1f} | "
                   f"{100 * sum(v) / tot:.1f}% |")
    return "\n".join(out)


def stats_table(path, top=20):
    rows = list(csv.DictReader(open(path)))
    out = ["| kernel | calls | avg us | total ms | share |", "|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append(f"| `{short_name(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.3f} | {float(r['Percentage']):.1f}% |")
    return "\n".join(out)


def counter_means(path, counter):
    d = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        d[(short_name(r["Kernel_Name"]), grid_of(r))].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in d.items()}, {k: len(v) for k, v in d.items()}


def pmc_report(fetch_csv, write_csv, copy_bytes):
    fetch, nf = counter_means(fetch_csv, "FETCH_SIZE")
    write, _ = counter_means(write_csv, "WRITE_SIZE")
    copies = [k for k in fetch if COPY_PATTERN.search(k[0])]
    if not copies:
        raise SystemExit("no copy kernel in the FETCH_SIZE run: cannot calibrate")
    ck = max(copies, key=lambda k: fetch[k])
    cf = copy_bytes / (fetch[ck] * 1024.0)
    cw = copy_bytes / (write[ck] * 1024.0) if ck in write else float("nan")
    lines = [f"Calibration: `{ck[0]}` copying {copy_bytes} B reads FETCH_SIZE {fetch[ck]:.0f} KiB and "
             f"WRITE_SIZE {write.get(ck, float('nan')):.0f} KiB per dispatch, so FETCH_SIZE is scaled by "
             f"{cf:.3f} and WRITE_SIZE by {cw:.3f}.", "",
             "| kernel | grid (threads) | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | calibrated HBM bytes |",
             "|---|---|---|---|---|---|"]
    level0 = {}
    for k in sorted(fetch, key=lambda k: -fetch[k]):
        if k not in write:
            continue
        b = fetch[k] * 1024.0 * cf + write[k] * 1024.0 * cw
        lines.append(f"| `{k[0]}` | {k[1]} | {nf[k]} | {fetch[k]:.0f} | {write[k]:.0f} | {b:.4g} |")
        _, prec = base_and_precision(k[0])
        cls = kernel_class(k[0])
        if cls not in LEVEL0_CLASSES:
            continue
        tag = f"{cls}@level0/{prec}"
        if tag not in level0 or b > level0[tag][1]:
            level0[tag] = (k[1], b)
    lines.append("")
    lines.append("Level-0 traffic per launch (the most traffic of each kernel class):")
    lines.append("")
    lines.append("```")
    for tag, (_, b) in sorted(level0.items()):
        lines.append(f"TRAFFIC {tag} {b:.0f}")
    lines.append("```")
    return "\n".join(lines)


def dispatch_values(path, counter):
    """[(dispatch id, short kernel name, grid, value)] in dispatch order
    (values of one dispatch summed over the rows that carry them)."""
    acc = collections.OrderedDict()
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        did = int(r.get("Dispatch_Id") or r.get("Correlation_Id") or len(acc))
        if did not in acc:
            acc[did] = [short_name(r["Kernel_Name"]), grid_of(r), 0.0]
        acc[did][2] += float(r["Counter_Value"])
    return [(d, v[0], v[1], v[2]) for d, v in sorted(acc.items())]


def last_solve(rows):
    """Dispatches after the last device copy (the restore before the solve)."""
    last = max((k for k, r in enumerate(rows) if COPY_PATTERN.search(r[1])), default=None)
    if last is None:
        raise SystemExit("no device copy in the run: cannot find the solve")
    return rows[last + 1:]


def solve_report(fetch_csv, write_csv, copy_bytes, label):
    fetch = dispatch_values(fetch_csv, "FETCH_SIZE")
    write = dispatch_values(write_csv, "WRITE_SIZE")
    cf_copy = [r[3] for r in fetch if COPY_PATTERN.search(r[1])]
    cw_copy = [r[3] for r in write if COPY_PATTERN.search(r[1])]
    cf = copy_bytes / (max(cf_copy) * 1024.0)
    cw = copy_bytes / (max(cw_copy) * 1024.0)
    fs, ws = last_solve(fetch), last_solve(write)
    if [(r[1], r[2]) for r in fs] != [(r[1], r[2]) for r in ws]:
        raise SystemExit("the FETCH_SIZE and WRITE_SIZE runs dispatched different solves")
    per = collections.defaultdict(lambda: [0, 0.0])
    total = 0.0
    for f, w in zip(fs, ws):
        b = f[3] * 1024.0 * cf + w[3] * 1024.0 * cw
        per[(f[1], f[2])][0] += 1
        per[(f[1], f[2])][1] += b
        total += b
    lines = [f"Calibration: the run's device copy of {copy_bytes:.0f} B reads FETCH_SIZE {max(cf_copy):.0f} KiB and "
             f"WRITE_SIZE {max(cw_copy):.0f} KiB, so FETCH_SIZE is scaled by {cf:.3f} and WRITE_SIZE by {cw:.3f}.",
             "", f"One solve ({label}): {len(fs)} dispatches, {total:.4g} calibrated HBM bytes.", "",
             "| kernel | grid (threads) | dispatches | calibrated HBM bytes | share |", "|---|---|---|---|---|"]
    for (name, grid), (n, b) in sorted(per.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| `{name}` | {grid} | {n} | {b:.4g} | {100 * b / total:.1f}% |")
    lines += ["", "```", f"SOLVE_TRAFFIC {label.replace(' ', '')} {total:.0f} {len(fs)}", "```"]
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="mode", required=True)
    t = sub.add_parser("trace")
    t.add_argument("kernel_trace")
    t.add_argument("kernel_stats", nargs="?")
    p = sub.add_parser("pmc")
    p.add_argument("fetch_csv")
    p.add_argument("write_csv")
    p.add_argument("--copy-bytes", type=float, required=True)
    q = sub.add_parser("solve")
    q.add_argument("fetch_csv")
    q.add_argument("write_csv")
    q.add_argument("--copy-bytes", type=float, required=True)
    q.add_argument("--label", required=True)
    a = ap.parse_args(argv)
    if a.mode == "trace":
        print(trace_table(a.kernel_trace))
        if a.kernel_stats:
            print()
            print(stats_table(a.kernel_stats))
    elif a.mode == "pmc":
        print(pmc_report(a.fetch_csv, a.write_csv, a.copy_bytes))
    else:
        print(solve_report(a.fetch_csv, a.write_csv, a.copy_bytes, a.label))


if __name__ == "__main__":
    sys.exit(main())
