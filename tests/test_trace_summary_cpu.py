"""das_amd.trace_summary: the union of overlapping kernel intervals, and the
table it prints from a rocprofv3 kernel-trace CSV."""
import csv

from das_amd import trace_summary as ts


def test_union_of_intervals():
    assert ts.union_ns([]) == 0
    assert ts.union_ns([(0, 10)]) == 10
    # disjoint, nested, touching and overlapping intervals, in any order
    assert ts.union_ns([(20, 30), (0, 10)]) == 20
    assert ts.union_ns([(0, 100), (10, 20), (30, 40)]) == 100
    assert ts.union_ns([(0, 10), (10, 20)]) == 20
    assert ts.union_ns([(5, 15), (0, 10), (12, 30), (40, 41)]) == 31


def test_summary_of_a_trace(tmp_path, capsys):
    p = tmp_path / "t_kernel_trace.csv"
    rows = [("KERNEL_DISPATCH", "k_tc<5, 6>(...)", 0, 100), ("KERNEL_DISPATCH", "k_tc<5, 6>(...)", 50, 150),
            ("KERNEL_DISPATCH", "k_tc<5, 6>(...)", 300, 400), ("KERNEL_DISPATCH", "k_fill", 150, 160),
            ("MEMORY_COPY", "k_fill", 0, 1000)]
    with open(p, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kind", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writerows(rows)
    out = {r["kernel"]: r for r in ts.summarize(str(p))}
    assert out["k_tc<5, 6>(...)"]["calls"] == 3
    assert out["k_tc<5, 6>(...)"]["sum_ms"] == 300 / 1e6
    assert out["k_tc<5, 6>(...)"]["union_ms"] == 250 / 1e6
    assert out["k_fill"]["calls"] == 1  # the copy record is not a dispatch
    import sys

    argv = sys.argv
    sys.argv = ["trace_summary", str(p), "--match", "k_tc", "--launches", "2"]
    try:
        ts.main()
    finally:
        sys.argv = argv
    text = capsys.readouterr().out
    assert "union 0.00 ms" in text and "0.12 us per launch over 2 launches" in text
