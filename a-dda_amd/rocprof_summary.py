"""Condense rocprofv3 CSV output into the short per-kernel text summaries kept with the benchmark results.

    python a-dda_amd/rocprof_summary.py stats DIR        # from `rocprofv3 --kernel-trace --stats --output-format csv`
    python a-dda_amd/rocprof_summary.py pmc DIR COUNTER  # from `rocprofv3 --pmc COUNTER --output-format csv`

`stats` prints, per kernel (template arguments kept, parameter list dropped): calls, average / min / max duration
in microseconds and total milliseconds, from the *_kernel_stats.csv (or *_kernel_trace.csv) found under DIR.
`pmc` prints, per kernel and grid size, the number of dispatches and the average counter value per dispatch (the
FETCH_SIZE / WRITE_SIZE counters are in KiB), from the *_counter_collection.csv under DIR.
"""
import csv
import glob
import os
import sys


def _find(d, suffix):
    hits = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not hits:
        raise SystemExit(f"no *{suffix} under {d}")
    return hits


def _col(row, *names):
    low = {k.lower(): v for k, v in row.items()}
    for n in names:
        if n.lower() in low:
            return low[n.lower()]
    raise KeyError(f"none of {names} in CSV columns {list(row)}")


def kname(s):
    s = s.strip()
    depth = 0
    for i, ch in enumerate(s):  # cut the parameter list: first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            s = s[:i]
            break
    return s[5:] if s.startswith("void ") else s


def stats(d):
    files = glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)
    agg = {}
    if files:
        for f in files:
            for r in csv.DictReader(open(f)):
                dur = (int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))) / 1e3
                a = agg.setdefault(kname(_col(r, "Kernel_Name")), [])
                a.append(dur)
        rows = [(k, len(v), sum(v) / len(v), min(v), max(v), sum(v) / 1e3) for k, v in agg.items()]
    else:
        rows = []
        for f in _find(d, "kernel_stats.csv"):
            for r in csv.DictReader(open(f)):
                rows.append((kname(_col(r, "Name")), int(_col(r, "Calls")), float(_col(r, "AverageNs")) / 1e3,
                             float(_col(r, "MinNs")) / 1e3, float(_col(r, "MaxNs")) / 1e3,
                             float(_col(r, "TotalDurationNs")) / 1e6))
    rows.sort(key=lambda t: -t[5])
    print(f"{'kernel':<58} {'calls':>6} {'avg_us':>9} {'min_us':>9} {'max_us':>9} {'total_ms':>10}")
    for k, n, avg, mn, mx, tot in rows:
        print(f"{k[:58]:<58} {n:>6} {avg:>9.1f} {mn:>9.1f} {mx:>9.1f} {tot:>10.3f}")


def pmc(d, counter):
    agg = {}
    for f in _find(d, "counter_collection.csv"):
        for r in csv.DictReader(open(f)):
            if not _col(r, "Counter_Name").startswith(counter):
                continue
            key = (kname(_col(r, "Kernel_Name")), int(_col(r, "Grid_Size")))
            agg.setdefault(key, []).append(float(_col(r, "Counter_Value")))
    print(f"{'kernel':<58} {'grid':>10} {'n':>5} {counter + ' avg/dispatch':>24}")
    for (k, g), v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        print(f"{k[:58]:<58} {g:>10} {len(v):>5} {sum(v) / len(v):>24.1f}")


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "stats":
        stats(sys.argv[2])
    elif len(sys.argv) >= 4 and sys.argv[1] == "pmc":
        pmc(sys.argv[2], sys.argv[3])
    else:
        raise SystemExit(__doc__)
