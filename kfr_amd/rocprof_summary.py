"""Markdown summaries of rocprofv3 result databases (kernel trace and PMC counters).

    python -m kfr_amd.rocprof_summary trace  <dir>/run_results.db [--top 40]
    python -m kfr_amd.rocprof_summary counter <dir>/run_results.db FETCH_SIZE [--scale 2.0] [--match k_pow2]

`trace` lists kernels by total time (calls, total, average, share). `counter` lists, per kernel, the
average value of one counter per dispatch (KiB as rocprofv3 reports it, and MiB after multiplying
by --scale, the factor that makes the counter read a plain device copy of known size correctly).
"""
from __future__ import annotations

import argparse
import sqlite3
import sys
from collections import OrderedDict


def _short(name: str, width: int = 100) -> str:
    name = name.replace("kfr_amd::", "").replace("p2::", "")
    return name if len(name) <= width else name[: width - 3] + "..."


def trace(db: str, top: int) -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select name, total_calls, total_duration, average, percentage from top_kernels "
                     "order by total_duration desc").fetchall()
    # top_kernels reports microseconds
    out = ["| calls | total ms | avg us | share % | kernel |", "|---:|---:|---:|---:|---|"]
    for name, calls, tot, avg, pct in rows[:top]:
        out.append(f"| {calls} | {tot / 1e3:.3f} | {avg:.1f} | {pct:.1f} | `{_short(name)}` |")
    return "\n".join(out)


def counter(db: str, cname: str, scale: float, match: str | None) -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, value from counters_collection where counter_name = ? "
                     "order by dispatch_id", (cname,)).fetchall()
    agg: "OrderedDict[str, list[float]]" = OrderedDict()
    for name, v in rows:
        if match and match not in name:
            continue
        agg.setdefault(name, []).append(float(v))
    out = [f"| dispatches | {cname} KiB/dispatch (avg) | x{scale:g} = MiB/dispatch (avg) | largest dispatch, MiB "
           f"(x{scale:g}) | kernel |", "|---:|---:|---:|---:|---|"]
    for name, vs in agg.items():
        avg = sum(vs) / len(vs)
        out.append(f"| {len(vs)} | {avg:.1f} | {avg * scale / 1024:.1f} | {max(vs) * scale / 1024:.1f} | "
                   f"`{_short(name)}` |")
    return "\n".join(out)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    t.add_argument("--top", type=int, default=40)
    k = sub.add_parser("counter")
    k.add_argument("db")
    k.add_argument("name")
    k.add_argument("--scale", type=float, default=1.0)
    k.add_argument("--match", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        print(trace(a.db, a.top))
    else:
        print(counter(a.db, a.name, a.scale, a.match))
    return 0


if __name__ == "__main__":
    sys.exit(main())
