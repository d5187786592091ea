"""Condense rocprofv3 CSV output into the short per-kernel text summaries kept with the benchmark results.

    python a-dda_amd/rocprof_summary.py stats DIR        # from `rocprofv3 --kernel-trace --stats --output-format csv`
    python a-dda_amd/rocprof_summary.py pmc DIR COUNTER  # from `rocprofv3 --pmc COUNTER --output-format csv`
    python a-dda_amd/rocprof_summary.py traffic FETCHDIR WRITEDIR BENCH_JSON COPY_BYTES  # HBM bytes per pass kernel
    python a-dda_amd/rocprof_summary.py sq SQDIR TRACEDIR  # per-wave instruction mix and VALU issue time
    python a-dda_amd/rocprof_summary.py groups DIR NITER   # a solve's kernel time by group, per iteration

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


def _pmc_avg(d, counter):
    """{(kernel, grid): average counter value per dispatch} of a --pmc run."""
    agg = {}
    for f in _find(d, "counter_collection.csv"):
        for r in csv.DictReader(open(f)):
            if _col(r, "Counter_Name").startswith(counter):
                key = (kname(_col(r, "Kernel_Name")), int(_col(r, "Grid_Size")))
                agg.setdefault(key, []).append(float(_col(r, "Counter_Value")))
    return {k: sum(v) / len(v) for k, v in agg.items()}


def _trace_avg(d):
    """{kernel: average duration in us} of a --kernel-trace run."""
    out = {}
    for f in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            k = kname(_col(r, "Kernel_Name"))
            out.setdefault(k, []).append((int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))) / 1e3)
    return {k: sum(v) / len(v) for k, v in out.items()}


def _pass_of(kernel):
    """(pass index, precision) of a matvec pass kernel name (P1 k_p1/k_p1s, P2/P4 k_zfft forward/inverse, P3 k_p3*,
    P5 k_p5/k_p5s), or None."""
    t = "double" if "<double" in kernel else "float" if "<float" in kernel else None
    base = kernel.split("<")[0].split("::")[-1]
    if t is None:
        return None
    if base in ("k_p1", "k_p1s"):
        return 0, t
    if base == "k_zfft":
        return (3 if f"<{t}, true" in kernel else 1), t
    if base.startswith("k_p3"):
        return 2, t
    if base in ("k_p5", "k_p5s"):
        return 4, t
    return None


def traffic(fetch_dir, write_dir, bench_json, copy_bytes):
    """HBM bytes per dispatch of the matvec pass kernels from FETCH_SIZE / WRITE_SIZE (KiB), each scaled by what the
    counter reads for the plain copy of copy_bytes bytes (__amd_rocclr_copyBuffer, the bench's calibration copy),
    against the algorithmic bytes of the bench's --matvec-only JSON line."""
    import json
    fe, wr = _pmc_avg(fetch_dir, "FETCH_SIZE"), _pmc_avg(write_dir, "WRITE_SIZE")
    kib = copy_bytes / 1024
    cf = [v for (k, g), v in fe.items() if k == "__amd_rocclr_copyBuffer" and abs(v - kib / 2) < 0.05 * kib]
    cw = [v for (k, g), v in wr.items() if k == "__amd_rocclr_copyBuffer" and abs(v - kib) < 0.05 * kib]
    sf, sw = kib / cf[0], kib / cw[0]
    mv = json.load(open(bench_json))["matvec"]
    print(f"Calibration: the copy of {copy_bytes} B ({kib:.1f} KiB) reads as FETCH {cf[0]:.1f} KiB and WRITE "
          f"{cw[0]:.1f} KiB: read bytes = FETCH x {sf:.4f} x 1024, write bytes = WRITE x {sw:.4f} x 1024.\n")
    print(f"{'pass':<5}{'kernel':<52}{'FETCH KiB':>12}{'WRITE KiB':>12}{'HBM bytes':>13}{'algorithmic':>13}{'ratio':>7}")
    rows = []
    for (k, g), f in fe.items():
        pt = _pass_of(k)
        if pt is None or (k, g) not in wr:
            continue
        i, t = pt
        alg = mv["fp64" if t == "double" else "fp32"]["pass_bytes"][i]
        hbm = (f * sf + wr[(k, g)] * sw) * 1024
        rows.append((t != "double", i, k, f, wr[(k, g)], hbm, alg))
    for _, i, k, f, w, hbm, alg in sorted(rows):
        print(f"P{i + 1:<4}{k[:51]:<52}{f:>12.1f}{w:>12.1f}{hbm:>13.4e}{alg:>13.4e}{hbm / alg:>7.2f}")


_VECTOR = ("k_dk_lc", "k_dk_xr", "k_dk_qmr_upd", "k_dot", "k_lincomb", "k_convert")
_BUILD = ("k_green", "k_rowfft", "k_colfft", "k_extract_d", "k_green_refl", "k_extract_r")


def _group_of(kernel):
    base = kernel.split("<")[0].split("::")[-1]
    pt = _pass_of(kernel)
    if pt is not None:
        return f"P{pt[0] + 1} {pt[1]}"
    if base in _VECTOR:
        return "vector kernels " + ("double" if "<double" in kernel else "float")
    if base.startswith("k_dk_") or base == "k_reduce_final":
        return "scalar kernels"
    if base in _BUILD:
        return "D-matrix build"
    return "other"


def groups(d, niter):
    """Kernel time of a solve grouped into matvec passes (per precision), Krylov vector kernels, the one-block scalar
    kernels and the rest, in total and per iteration (niter iterations)."""
    agg = {}
    for f in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        for r in csv.DictReader(open(f)):
            g = _group_of(kname(_col(r, "Kernel_Name")))
            t = (int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))) / 1e3
            a = agg.setdefault(g, [0, 0.0])
            a[0] += 1
            a[1] += t
    print(f"{'group':<24}{'calls':>8}{'total_ms':>11}{'us/iter':>10}")
    tot = 0.0
    for g in sorted(agg):
        n, t = agg[g]
        tot += t if g != "D-matrix build" else 0.0
        print(f"{g:<24}{n:>8}{t / 1e3:>11.3f}{t / niter:>10.1f}")
    print(f"{'all but D-matrix build':<24}{'':>8}{tot / 1e3:>11.3f}{tot / niter:>10.1f}")


SQ = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR")


def sq(sq_dir, trace_dir, ncu=256, ghz=2.4):
    """Per-wave instruction counts of the matvec pass kernels and their VALU issue time (a CU issues at most one wave64
    VALU instruction per cycle) against the kernel-trace duration."""
    c = {n: _pmc_avg(sq_dir, n) for n in SQ}
    dur = _trace_avg(trace_dir)
    print(f"{'kernel':<52}{'waves':>8}{'VALU/w':>8}{'SALU/w':>8}{'LDS/w':>7}{'RD/w':>6}{'WR/w':>6}{'valu us':>9}"
          f"{'dur us':>8}{'valu/dur':>9}")
    keys = sorted({kg for kg in c["SQ_WAVES"] if _pass_of(kg[0])}, key=lambda kg: (_pass_of(kg[0])[1] == "float",
                                                                                  _pass_of(kg[0])[0]))
    for kg in keys:
        w = c["SQ_WAVES"][kg]
        if not w:
            continue
        per = [c[n].get(kg, 0) / w for n in SQ[1:]]
        valu_us = c["SQ_INSTS_VALU"].get(kg, 0) / ncu / (ghz * 1e3)
        d = dur.get(kg[0], float("nan"))
        print(f"{kg[0][:51]:<52}{w:>8.0f}{per[0]:>8.0f}{per[1]:>8.0f}{per[2]:>7.0f}{per[3]:>6.1f}{per[4]:>6.1f}"
              f"{valu_us:>9.1f}{d:>8.1f}{valu_us / d:>9.2f}")


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "stats":
        stats(sys.argv[2])
    elif len(sys.argv) >= 4 and sys.argv[1] == "pmc":
        pmc(sys.argv[2], sys.argv[3])
    elif len(sys.argv) >= 6 and sys.argv[1] == "traffic":
        traffic(sys.argv[2], sys.argv[3], sys.argv[4], int(sys.argv[5]))
    elif len(sys.argv) >= 4 and sys.argv[1] == "groups":
        groups(sys.argv[2], int(sys.argv[3]))
    elif len(sys.argv) >= 4 and sys.argv[1] == "sq":
        sq(sys.argv[2], sys.argv[3])
    else:
        raise SystemExit(__doc__)
