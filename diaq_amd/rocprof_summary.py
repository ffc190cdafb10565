"""Summarise rocprofv3 results (the rocpd SQLite files) as short text: per-kernel dispatch count and average duration,
and per-dispatch FETCH_SIZE / WRITE_SIZE in bytes, each also scaled by what the same counter reads for a plain copy
kernel of known size (the calibration dispatch that ``bench.py --profile`` runs after its SpMV steps).

    python -m diaq_amd.rocprof_summary --stats DB [--fetch DB] [--write DB] [--match spmv_rows_kernel] [--json OUT]
"""
from __future__ import annotations

import argparse
import sqlite3
import statistics

COPY_BYTES = 1 << 31  # bench.py --profile: dst.copy_(src) of 2^28 doubles


def kernel_stats(db, match=None):
    c = sqlite3.connect(db)
    out = []
    for name, calls, total, avg, pct in c.execute("select name, total_calls, total_duration, average, percentage "
                                                  "from top_kernels"):
        if match and match not in name:
            continue
        out.append((name, calls, total, avg, pct))
    return out


def counter(db, cname, match):
    """Per-dispatch values (bytes) of one counter for the kernel matching `match` -- when several kernel names match
    (an autotuner's candidates), the one dispatched most often, i.e. the schedule the timed run used -- and the
    value the counter reads for the calibration copy."""
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, value, duration from counters_collection where counter_name=? order by id",
                     (cname,)).fetchall()
    by_name = {}
    for k, v, d in rows:
        if match in k:
            by_name.setdefault(k, []).append((v * 1024.0, d))
    name = max(by_name, key=lambda k: len(by_name[k])) if by_name else None
    hit = by_name.get(name, [])
    # calibration: the largest copy-like dispatch (the 2 GiB copy that the profiling run ends with)
    cal = [v * 1024.0 for k, v, d in rows if ("copyBuffer" in k or "elementwise" in k) and v * 1024.0 > 1e9]
    return hit, (max(cal) if cal else None), name


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--match", default="spmv_rows_kernel")
    ap.add_argument("--top", type=int, default=6)
    ap.add_argument("--json", help="also write {kernel, fetch, write, factors, traffic_bytes} to this file")
    a = ap.parse_args()
    summary = {}
    print("kernel stats (rocprofv3 --kernel-trace --stats), durations in microseconds")
    for name, calls, total, avg, pct in kernel_stats(a.stats)[: a.top]:
        print(f"  {avg:10.1f} us avg  {calls:4d} calls  {pct:5.1f}%  {name[:110]}")
    for label, db, cname in (("FETCH_SIZE", a.fetch, "FETCH_SIZE"), ("WRITE_SIZE", a.write, "WRITE_SIZE")):
        if not db:
            continue
        hit, cal, name = counter(db, cname, a.match)
        if not hit:
            print(f"{label}: no dispatch matching {a.match}")
            continue
        med = statistics.median(v for v, _ in hit)
        short = name.split("(")[0].replace("void ", "")
        line = f"{label} per dispatch of {short}: median {med / 1e9:.3f} GB over {len(hit)} dispatches"
        factor = COPY_BYTES / cal if cal else None
        if cal:
            line += (f"; the 2 GiB calibration copy reads {cal / 1e9:.3f} GB, factor {factor:.3f} -> "
                     f"{med * factor / 1e9:.3f} GB")
        print(line)
        summary[label] = dict(raw_bytes=med, factor=factor, bytes=med * (factor or 1.0), kernel=short)
    if a.json and "FETCH_SIZE" in summary and "WRITE_SIZE" in summary:
        import json
        summary["kernel"] = summary["FETCH_SIZE"]["kernel"]
        summary["traffic_bytes"] = summary["FETCH_SIZE"]["bytes"] + summary["WRITE_SIZE"]["bytes"]
        with open(a.json, "w") as f:
            json.dump(summary, f, indent=1)


if __name__ == "__main__":
    main()
