"""Summary of a rocprofv3 kernel trace (``--kernel-trace --output-format csv``):
per kernel name the dispatch count, the average and summed duration, and the
time during which at least one of its dispatches ran (the union of their
intervals).  Kernels issued on several streams at once overlap, so their
summed durations exceed the time they occupy the GPU; the union is that time.

    python -m das_amd.trace_summary <dir>/<name>_kernel_trace.csv \\
        [--match k_tc] [--launches N]

With ``--match`` and ``--launches`` it also prints the union of the matching
kernels' intervals divided by N: the time per launch of a batched run whose
launches are split into one kernel per launch group.
"""
import argparse
import csv
from collections import defaultdict


def union_ns(intervals):
    """Total length of the union of [start, end) intervals."""
    total, cur_s, cur_e = 0, None, None
    for s, e in sorted(intervals):
        if cur_e is None or s > cur_e:
            if cur_e is not None:
                total += cur_e - cur_s
            cur_s, cur_e = s, e
        elif e > cur_e:
            cur_e = e
    if cur_e is not None:
        total += cur_e - cur_s
    return total


def summarize(path):
    by_name = defaultdict(list)
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            if row.get("Kind", "KERNEL_DISPATCH") != "KERNEL_DISPATCH":
                continue
            s, e = int(row["Start_Timestamp"]), int(row["End_Timestamp"])
            by_name[row["Kernel_Name"]].append((s, e))
    out = []
    for name, iv in by_name.items():
        durs = [e - s for s, e in iv]
        out.append({"kernel": name, "calls": len(iv), "avg_us": sum(durs) / len(iv) / 1e3,
                    "sum_ms": sum(durs) / 1e6, "union_ms": union_ns(iv) / 1e6, "iv": iv})
    out.sort(key=lambda r: -r["sum_ms"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("csv")
    ap.add_argument("--match", default=None, help="substring of the main kernel's name")
    ap.add_argument("--launches", type=int, default=0, help="launches the main kernel's dispatches make up")
    args = ap.parse_args()
    rows = summarize(args.csv)
    total = sum(r["sum_ms"] for r in rows)
    print("| kernel | dispatches | average (us) | summed (ms) | union (ms) | share of summed |")
    print("|---|---|---|---|---|---|")
    for r in rows:
        name = r["kernel"] if len(r["kernel"]) < 60 else r["kernel"][:57] + "..."
        print(f"| `{name}` | {r['calls']} | {r['avg_us']:.2f} | {r['sum_ms']:.2f} | {r['union_ms']:.2f} | "
              f"{100 * r['sum_ms'] / total:.2f}% |")
    if args.match:
        iv = [x for r in rows if args.match in r["kernel"] for x in r["iv"]]
        u = union_ns(iv) / 1e6
        line = f"\n`{args.match}`: {len(iv)} dispatches, union {u:.2f} ms"
        if args.launches:
            line += f", {1e3 * u / args.launches:.2f} us per launch over {args.launches} launches"
        print(line)


if __name__ == "__main__":
    main()
