"""Per-kernel statistics from a rocprofv3 kernel trace (the ``*_results.db``
database, or the ``*_kernel_trace.csv`` of ``--output-format csv``): calls,
total and average duration per kernel name, largest total first, as a
Markdown table.

    python -m atomx_amd.native_bench.kernel_stats TRACE [--match TEXT] [--top N]
                                                  [--last N] [--per N] [--between K]
                                                  [--until K] [--after K] [--by-grid]

--last N keeps only the last N dispatches of each kernel (a timed loop that
follows a tuning sweep of the same kernels); --per N adds a column with the
total divided by N (time per SCF sweep, per request, ...); --between K keeps
the dispatches from the first to the last dispatch of kernel K; --until K
keeps those before the first dispatch of K, --after K those from it on
(phases of one traced run); --by-grid lists (kernel, grid x, grid y) groups
with their average and minimum duration instead (CSV traces: the launch
shape tells the schedules of one kernel apart).
"""

from __future__ import annotations

import argparse
import csv
import re
import sqlite3


def _short(name):
    s = name.replace("(anonymous namespace)::", "")
    s = re.sub(r"^void ", "", s)
    return re.sub(r"\(.*", "", s)


def dispatches(path):
    """[(start ns, short kernel name, duration ns)] of every kernel dispatch."""
    if path.endswith(".csv"):
        out = []
        for r in csv.DictReader(open(path)):
            t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            out.append((t0, _short(r["Kernel_Name"]), t1 - t0))
        return out
    con = sqlite3.connect(path)
    rows = con.execute("select start, name, duration from kernels").fetchall()
    con.close()
    return [(int(t0), _short(name), int(d)) for t0, name, d in rows]


def grid_groups(path, match=None, between=None, until=None, after=None):
    """{(kernel, grid x, grid y): [durations]} of a CSV kernel trace, in the
    same phase windows as kernel_stats."""
    disp = []
    for r in csv.DictReader(open(path)):
        t0 = int(r["Start_Timestamp"])
        disp.append((t0, _short(r["Kernel_Name"]), int(r["End_Timestamp"]) - t0,
                     int(r["Grid_Size_X"]), int(r["Grid_Size_Y"])))
    disp = _window(sorted(disp), between, until, after)
    out = {}
    for _, name, dur, gx, gy in disp:
        if match and match not in name:
            continue
        out.setdefault((name, gx, gy), []).append(dur)
    return out


def _marks(disp, kernel):
    marks = [x[0] for x in disp if x[1] == kernel]
    if not marks:
        raise SystemExit(f"no dispatch of {kernel}")
    return marks


def _window(disp, between, until, after):
    """Dispatches (sorted by start) inside the phase window."""
    if between:  # the window from the first to the last dispatch of that kernel
        marks = _marks(disp, between)
        disp = [x for x in disp if marks[0] <= x[0] <= marks[-1]]
    if until:
        first = _marks(disp, until)[0]
        disp = [x for x in disp if x[0] < first]
    if after:
        first = _marks(disp, after)[0]
        disp = [x for x in disp if x[0] >= first]
    return disp


def kernel_stats(path, match=None, last=0, between=None, until=None, after=None):
    disp = _window(sorted(dispatches(path)), between, until, after)
    by_name = {}
    for t0, name, dur in disp:
        if match and match not in name:
            continue
        by_name.setdefault(name, []).append(dur)
    rows = []
    for name, durs in by_name.items():
        if last:
            durs = durs[-last:]
        rows.append((name, len(durs), float(sum(durs))))
    return sorted(rows, key=lambda r: -r[2])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--match", default=None)
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--last", type=int, default=0)
    ap.add_argument("--per", type=float, default=0.0)
    ap.add_argument("--between", default=None,
                    help="only dispatches from the first to the last one of this kernel")
    ap.add_argument("--until", default=None,
                    help="only dispatches before the first one of this kernel")
    ap.add_argument("--after", default=None,
                    help="only dispatches from the first one of this kernel on")
    ap.add_argument("--by-grid", action="store_true",
                    help="(kernel, grid x, grid y) groups with average and minimum (CSV traces)")
    a = ap.parse_args()
    if a.by_grid:
        print("| kernel | grid x | grid y | dispatches | avg us | min us |")
        print("|---|---:|---:|---:|---:|---:|")
        groups = grid_groups(a.trace, a.match, a.between, a.until, a.after)
        for (name, gx, gy), durs in sorted(groups.items()):
            durs = durs[-a.last:] if a.last else durs
            print(f"| `{name}` | {gx} | {gy} | {len(durs)} | {sum(durs) / len(durs) / 1e3:.2f} | "
                  f"{min(durs) / 1e3:.2f} |")
        return
    rows = kernel_stats(a.trace, a.match, a.last, a.between, a.until, a.after)
    total = sum(r[2] for r in rows)
    per = f" | us per 1/{a.per:g}" if a.per else ""
    print(f"| kernel | calls | total us | avg us | share{per} |")
    print("|---|---:|---:|---:|---:|" + ("---:|" if a.per else ""))
    for name, n, tot in rows[:a.top]:
        extra = f" | {tot / 1e3 / a.per:.1f}" if a.per else ""
        print(f"| `{name}` | {n} | {tot / 1e3:.1f} | {tot / n / 1e3:.2f} | "
              f"{100 * tot / total:.1f} %{extra} |")
    print(f"\ntotal kernel time {total / 1e6:.3f} ms over {sum(r[1] for r in rows)} dispatches")


if __name__ == "__main__":
    main()
