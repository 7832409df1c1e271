"""Summaries of rocprofv3 output for the kernels of this library.

    python -m gallatin_amd.profiling stats <kernel_stats.csv> [--title T] [--command C]
        Markdown table of per-kernel launch counts and durations.

    python -m gallatin_amd.profiling traffic <fetch counter_collection.csv>
        <write counter_collection.csv> --copy-bytes B --workload W
        [--match key=int,...] [--out F]
        HBM bytes per dispatch from a FETCH_SIZE run and a WRITE_SIZE run.
        Each counter is scaled by what it reads for k_bench_copy moving B
        bytes (B read and B written) in the same run, so the result is in
        bytes actually moved, whatever unit the counter definition assumes.

    python -m gallatin_amd.profiling distribution <kernel_trace.csv> [--last N]
        [--match SUBSTRING]
        Markdown table of each kernel's launch-duration percentiles (p50,
        p90, p99, max) over its last N launches (all by default), for the
        kernels whose name contains SUBSTRING.

    python -m gallatin_amd.profiling counters <counter_collection.csv>...
        [--dispatches KERNEL]
        Markdown table of the mean of every counter per kernel over its
        dispatches, merging several PMC runs of the same program; with
        --dispatches, also each dispatch of KERNEL in launch order.
"""
import argparse
import collections
import csv
import json
import sys


def kernel_key(name):
    """`void k_churn_batched<256>(gal_state*, ...)` -> `k_churn_batched<256>`."""
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    name = name.replace("(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def stats_markdown(path, title=None, command=None):
    rows = list(csv.DictReader(open(path)))
    out = []
    if title:
        out.append(f"# {title}\n")
    if command:
        out.append("Command:\n\n```\n" + command + "\n```\n")
    out.append("| kernel | calls | avg ms | min ms | max ms | share |")
    out.append("|---|---|---|---|---|---|")
    for r in rows:
        out.append(f"| `{kernel_key(r['Name'])}` | {r['Calls']} | "
                   f"{float(r['AverageNs']) / 1e6:.3f} | {float(r['MinNs']) / 1e6:.3f} | "
                   f"{float(r['MaxNs']) / 1e6:.3f} | {float(r['Percentage']):.2f} % |")
    return "\n".join(out) + "\n"


def distribution_markdown(path, last=0, match=""):
    by = collections.OrderedDict()
    for r in csv.DictReader(open(path)):
        k = kernel_key(r["Kernel_Name"])
        if match and match not in k:
            continue
        by.setdefault(k, []).append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    out = ["| kernel | launches | p50 ms | p90 ms | p99 ms | max ms |",
           "|---|---|---|---|---|---|"]
    for k, v in by.items():
        if last:
            v = v[-last:]
        w = sorted(v)

        def q(f):
            return w[min(len(w) - 1, int(f * (len(w) - 1) + 0.5))]
        out.append(f"| `{k}` | {len(w)} | {q(0.5):.3f} | {q(0.9):.3f} | {q(0.99):.3f} | "
                   f"{w[-1]:.3f} |")
    return "\n".join(out) + "\n"


def _per_kernel(path, counter):
    vals = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] == counter:
            vals[kernel_key(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return vals


def traffic(fetch_csv, write_csv, copy_bytes, workload, match=None):
    fetch = _per_kernel(fetch_csv, "FETCH_SIZE")
    write = _per_kernel(write_csv, "WRITE_SIZE")
    if "k_bench_copy" not in fetch or "k_bench_copy" not in write:
        raise SystemExit("both runs must include k_bench_copy for the scale factors")
    # the first copy launch is a warm-up that may fault pages in; use the median
    fc = sorted(fetch["k_bench_copy"])[len(fetch["k_bench_copy"]) // 2]
    wc = sorted(write["k_bench_copy"])[len(write["k_bench_copy"]) // 2]
    f_scale = copy_bytes / (fc * 1024.0)
    w_scale = copy_bytes / (wc * 1024.0)
    out = {"workload": workload, "match": match or {}, "copy_bytes": copy_bytes,
           "fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    for k in sorted(set(fetch) & set(write)):
        if k.startswith("__amd") or k.startswith("at::") or k.startswith("void at::"):
            continue
        f = sum(fetch[k]) / len(fetch[k])
        w = sum(write[k]) / len(write[k])
        out["kernels"][k] = {
            "dispatches": len(fetch[k]),
            "fetch_kib": f, "write_kib": w,
            "bytes": f * 1024.0 * f_scale + w * 1024.0 * w_scale,
        }
    return out


def _counter_rows(paths):
    """{kernel: {counter: [values in dispatch order]}} over several runs."""
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in paths:
        rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r.get("Dispatch_Id", 0)))
        for r in rows:
            out[kernel_key(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return out


def counters_markdown(paths, dispatches=None):
    data = _counter_rows(paths)
    names = sorted({c for k in data.values() for c in k})
    out = ["| kernel | dispatches | " + " | ".join(f"`{n}`" for n in names) + " |",
           "|---|---|" + "---|" * len(names)]
    for k in sorted(data):
        if k.startswith("__amd") or k.startswith("at::"):
            continue
        cells = []
        for n in names:
            v = data[k].get(n)
            cells.append(f"{sum(v) / len(v):,.0f}" if v else "")
        n_disp = max(len(v) for v in data[k].values())
        out.append(f"| `{k}` | {n_disp} | " + " | ".join(cells) + " |")
    if dispatches and dispatches in data:
        out.append("")
        out.append(f"Dispatches of `{dispatches}` in launch order:")
        out.append("")
        out.append("| # | " + " | ".join(f"`{n}`" for n in names if n in data[dispatches]) + " |")
        cols = [n for n in names if n in data[dispatches]]
        out.append("|---|" + "---|" * len(cols))
        for i in range(max(len(data[dispatches][n]) for n in cols)):
            out.append(f"| {i} | " + " | ".join(
                f"{data[dispatches][n][i]:,.0f}" if i < len(data[dispatches][n]) else ""
                for n in cols) + " |")
    return "\n".join(out) + "\n"


def main(argv=None):
    p = argparse.ArgumentParser()
    sub = p.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--title")
    s.add_argument("--command")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--workload", required=True)
    t.add_argument("--match", default="",
                   help="key=int,... : the benchmark settings this holds for")
    t.add_argument("--out")
    d = sub.add_parser("distribution")
    d.add_argument("csv")
    d.add_argument("--last", type=int, default=0)
    d.add_argument("--match", default="")
    c = sub.add_parser("counters")
    c.add_argument("csv", nargs="+")
    c.add_argument("--dispatches")
    a = p.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(stats_markdown(a.csv, a.title, a.command))
    elif a.cmd == "distribution":
        sys.stdout.write(distribution_markdown(a.csv, a.last, a.match))
    elif a.cmd == "counters":
        sys.stdout.write(counters_markdown(a.csv, a.dispatches))
    else:
        match = dict((kv.split("=")[0], int(kv.split("=")[1]))
                     for kv in a.match.split(",") if kv)
        res = traffic(a.fetch_csv, a.write_csv, a.copy_bytes, a.workload, match)
        text = json.dumps(res, indent=1) + "\n"
        if a.out:
            open(a.out, "w").write(text)
        sys.stdout.write(text)


if __name__ == "__main__":
    main()
