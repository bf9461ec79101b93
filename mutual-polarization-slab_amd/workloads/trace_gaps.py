"""GPU idle time inside solves, from a rocprofv3 kernel trace (CSV) of
workloads.solve_loop: kernels are grouped into solves at host gaps longer
than --split ms (solve_loop separates the solves with synchronisations and
a 10 ms host pause);
per solve, the span from its first kernel start to its last kernel end, the
summed kernel time, the idle time between kernels (overlaps merged), and the
largest gaps with the kernels on either side.

    python -m workloads.trace_gaps <kernel_trace.csv> [--split 2.0] [--top 8]
"""
from __future__ import annotations

import argparse
import csv


def load(path):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    return rows


def short(name):
    name = name.split("(")[0]
    return name.replace("void ", "").replace("mpol::", "")


def solves(rows, split_ns):
    groups, cur = [], [rows[0]]
    for r in rows[1:]:
        if r[0] - max(e for _, e, _ in cur[-8:]) > split_ns:
            groups.append(cur)
            cur = []
        cur.append(r)
    groups.append(cur)
    return groups


def analyse(group, top):
    start = group[0][0]
    end_max = group[0][1]
    busy = group[0][1] - group[0][0]
    gaps = []
    for prev, r in zip(group, group[1:]):
        if r[0] > end_max:
            gaps.append((r[0] - end_max, short(prev[2]), short(r[2])))
            busy += r[1] - r[0]
        else:
            busy += max(0, r[1] - end_max)
        end_max = max(end_max, r[1])
    span = end_max - start
    gaps.sort(reverse=True)
    return span, busy, gaps[:top], sum(g for g, _, _ in gaps), len(group)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--split", type=float, default=2.0)
    ap.add_argument("--top", type=int, default=8)
    a = ap.parse_args(argv)
    rows = load(a.trace)
    for k, g in enumerate(solves(rows, a.split * 1e6)):
        span, busy, gaps, idle, nk = analyse(g, a.top)
        print(f"group {k}: {nk} kernels, span {span / 1e6:.3f} ms, kernels {busy / 1e6:.3f} ms, "
              f"idle {idle / 1e6:.3f} ms")
        for gap, before, after in gaps:
            print(f"    gap {gap / 1e3:8.1f} us  after {before}  before {after}")


if __name__ == "__main__":
    main()
