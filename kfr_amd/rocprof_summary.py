"""Markdown summaries of rocprofv3 result databases (kernel trace and PMC counters).

    python -m kfr_amd.rocprof_summary trace  <dir>/run_results.db [--top 40]
    python -m kfr_amd.rocprof_summary counter <dir>/run_results.db FETCH_SIZE [--scale 2.0] [--match k_pow2]
    python -m kfr_amd.rocprof_summary traffic <fetch>/run_results.db <write>/run_results.db
    python -m kfr_amd.rocprof_summary table <dir>/run_results.db MfmaUtil VALUBusy SQ_WAIT_ANY ... [--match k_pow2]

`trace` lists kernels by total time (calls, total, average, share). `counter` lists, per kernel, the
average value of one counter per dispatch (KiB as rocprofv3 reports it, and MiB after multiplying
by --scale, the factor that makes the counter read a plain device copy of known size correctly).
`traffic` combines a FETCH_SIZE run and a WRITE_SIZE run into HBM bytes per dispatch, with both scale
factors taken from the calibration copy in the runs themselves. `table` lists, per kernel, the average
per dispatch of several counters of one run, and the share of wave cycles spent waiting
(SQ_WAIT_ANY / SQ_WAVE_CYCLES) when both are in it.
"""
from __future__ import annotations

import argparse
import sqlite3
import sys
from collections import OrderedDict


def _short(name: str, width: int = 100) -> str:
    name = name.replace("kfr_amd::", "").replace("p2::", "")
    return name if len(name) <= width else name[: width - 3] + "..."


def kernel_key(name: str) -> str:
    """A kernel's template instance without return type, namespaces and parameter list:
    'void kfr_amd::p2::k_pow2<float, 12, false, 2, 0>(kfr_amd::p2::args<float>)' ->
    'k_pow2<float, 12, false, 2, 0>' (the spelling of the library's kernel timing records)."""
    n = name.strip()
    if n.startswith("void "):
        n = n[5:]
    for ns in ("(anonymous namespace)::", "kfr_amd::", "p2::", "mx::"):
        n = n.replace(ns, "")
    depth = 0
    for i, ch in enumerate(n):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and i > 0:
            n = n[:i]
            break
    return n


def _per_kernel(db: str, cname: str) -> "OrderedDict[str, list[float]]":
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, value from counters_collection where counter_name = ? "
                     "order by dispatch_id", (cname,)).fetchall()
    agg: "OrderedDict[str, list[float]]" = OrderedDict()
    for name, v in rows:
        agg.setdefault(name, []).append(float(v))
    return agg


def traffic(fetch_db: str, write_db: str, probe_match: str, probe_bytes: int) -> str:
    """HBM bytes per dispatch from a FETCH_SIZE run and a WRITE_SIZE run. Each counter is scaled by
    the factor that makes it read the calibration probe correctly: the largest dispatch of the kernel
    matching probe_match, a plain device copy that reads and writes probe_bytes."""
    f = _per_kernel(fetch_db, "FETCH_SIZE")
    w = _per_kernel(write_db, "WRITE_SIZE")

    def probe(agg, cname):
        vals = [v for name, vs in agg.items() if probe_match in name for v in vs]
        if not vals:
            raise SystemExit(f"no dispatch of {probe_match!r} in the {cname} run")
        return max(vals)

    pf, pw = probe(f, "FETCH_SIZE"), probe(w, "WRITE_SIZE")
    sf, sw = probe_bytes / (pf * 1024.0), probe_bytes / (pw * 1024.0)
    out = [f"Calibration: the probe ({probe_match}, {probe_bytes / 2**20:.0f} MiB read and written) reads "
           f"FETCH_SIZE = {pf:.0f} KiB and WRITE_SIZE = {pw:.0f} KiB: scale factors "
           f"FETCH x{sf:.3f}, WRITE x{sw:.3f}.", "",
           "| dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | traffic bytes/dispatch | kernel |",
           "|---:|---:|---:|---:|---|"]
    rows = []
    for name, fv in f.items():
        wv = w.get(name)
        if not wv or probe_match in name:
            continue
        fa, wa = sum(fv) / len(fv), sum(wv) / len(wv)
        rows.append((len(fv), fa, wa, (fa * sf + wa * sw) * 1024.0, kernel_key(name)))
    rows.sort(key=lambda r: -r[3] * r[0])
    for n, fa, wa, tb, key in rows:
        out.append(f"| {n} | {fa:.1f} | {wa:.1f} | {tb:.0f} | `{_short(key)}` |")
    return "\n".join(out)


def read_traffic(path: str, key: str) -> float | None:
    """traffic bytes per dispatch of kernel `key` from a table written by `traffic`, or None"""
    try:
        with open(path) as fh:
            for line in fh:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 5 and cells[4] == f"`{key}`":
                    return float(cells[3])
    except OSError:
        return None
    return None


def read_table(path: str) -> list[dict]:
    """rows of the first table written by `table` in a markdown file (column name -> cell text, the
    kernel key under 'kernel'), or [] when the file is absent"""
    rows: list[dict] = []
    head: list[str] | None = None
    try:
        with open(path) as fh:
            for line in fh:
                if not line.startswith("|"):
                    if head is not None and rows:
                        break
                    continue
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if head is None:
                    head = cells
                elif not set(cells[0]) <= set("-:"):
                    row = dict(zip(head, cells))
                    row["kernel"] = row.get("kernel", "").strip("`")
                    rows.append(row)
    except OSError:
        return []
    return rows


def trace(db: str, top: int) -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select name, total_calls, total_duration, average, percentage from top_kernels "
                     "order by total_duration desc").fetchall()
    # top_kernels reports microseconds
    out = ["| calls | total ms | avg us | share % | kernel |", "|---:|---:|---:|---:|---|"]
    for name, calls, tot, avg, pct in rows[:top]:
        out.append(f"| {calls} | {tot / 1e3:.3f} | {avg:.1f} | {pct:.1f} | `{_short(kernel_key(name))}` |")
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


def table(db: str, names: list[str], match: str | None) -> str:
    per = {n: _per_kernel(db, n) for n in names}
    kernels: "OrderedDict[str, None]" = OrderedDict()
    for n in names:
        for k in per[n]:
            if not match or match in k:
                kernels[k] = None
    wait = "SQ_WAIT_ANY" in names and "SQ_WAVE_CYCLES" in names
    head = "| dispatches | " + " | ".join(names) + (" | wait share |" if wait else " |") + " kernel |"
    out = [head, "|" + "---:|" * (len(names) + 1 + (1 if wait else 0)) + "---|"]
    for k in kernels:
        avg = {n: (sum(per[n][k]) / len(per[n][k]) if per[n].get(k) else float("nan")) for n in names}
        cnt = max(len(per[n].get(k, [])) for n in names)
        cells = " | ".join(f"{avg[n]:.4g}" for n in names)
        extra = f" | {avg['SQ_WAIT_ANY'] / avg['SQ_WAVE_CYCLES']:.3f}" if wait and avg["SQ_WAVE_CYCLES"] else ""
        out.append(f"| {cnt} | {cells}{extra} | `{_short(kernel_key(k))}` |")
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
    h = sub.add_parser("traffic")
    h.add_argument("fetch_db")
    h.add_argument("write_db")
    h.add_argument("--probe-match", default="__amd_rocclr_copyBuffer")
    h.add_argument("--probe-bytes", type=int, default=256 << 20)
    g = sub.add_parser("table")
    g.add_argument("db")
    g.add_argument("names", nargs="+")
    g.add_argument("--match", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "table":
        print(table(a.db, a.names, a.match))
    elif a.cmd == "trace":
        print(trace(a.db, a.top))
    elif a.cmd == "traffic":
        print(traffic(a.fetch_db, a.write_db, a.probe_match, a.probe_bytes))
    else:
        print(counter(a.db, a.name, a.scale, a.match))
    return 0


if __name__ == "__main__":
    sys.exit(main())
