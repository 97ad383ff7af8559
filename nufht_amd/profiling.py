"""Text summaries of rocprofv3 results (rocpd SQLite databases).

    python -m nufht_amd.profiling trace DB [--match REGEX]
        per-kernel calls, total / average / min / max duration (ms), share

    python -m nufht_amd.profiling counters DB [DB ...] [--match REGEX]
        per-kernel counter totals and per-dispatch averages

    python -m nufht_amd.profiling traffic FETCH_DB WRITE_DB --kernel NAME
            --copy REGEX --copy-bytes N
        HBM bytes per dispatch of NAME: the average FETCH_SIZE and WRITE_SIZE
        (KiB per dispatch, summed over the counter's instances) of NAME, each
        scaled by (N bytes) / (the same counter's largest reading for the
        streaming kernel matching REGEX, which reads N bytes and writes N bytes)

Used to produce the summaries kept with the sources; reads only the
databases it is given.
"""
from __future__ import annotations

import argparse
import collections
import re
import sqlite3
import sys
from typing import Dict, List


def kernel_durations(db: str) -> Dict[str, List[float]]:
    c = sqlite3.connect(db)
    cols = [r[1] for r in c.execute("pragma table_info(kernels)")]
    name = "kernel_name" if "kernel_name" in cols else "name"
    out: Dict[str, List[float]] = collections.OrderedDict()
    for n, dur in c.execute(f"select {name}, end - start from kernels order by start"):
        out.setdefault(n, []).append(dur / 1e6)
    return out


def trace_table(db: str, match: str = "") -> str:
    d = kernel_durations(db)
    tot = sum(sum(v) for v in d.values())
    rx = re.compile(match) if match else None
    lines = [f"{'kernel':64s} {'calls':>6s} {'total_ms':>10s} {'avg_ms':>9s} {'min_ms':>8s} "
             f"{'max_ms':>8s} {'pct':>6s}"]
    for k, v in sorted(d.items(), key=lambda kv: -sum(kv[1])):
        if rx and not rx.search(k):
            continue
        s = sum(v)
        lines.append(f"{k[:64]:64s} {len(v):6d} {s:10.3f} {s / len(v):9.4f} {min(v):8.4f} "
                     f"{max(v):8.4f} {100 * s / tot:6.1f}")
    return "\n".join(lines)


def counter_values(dbs: List[str]) -> Dict[str, Dict[str, Dict[int, float]]]:
    """kernel -> counter -> dispatch -> value (summed over counter instances)."""
    out: Dict[str, Dict[str, Dict[int, float]]] = {}
    for db in dbs:
        c = sqlite3.connect(db)
        q = "select dispatch_id, kernel_name, counter_name, value from counters_collection"
        for did, kn, cn, v in c.execute(q):
            dd = out.setdefault(kn, {}).setdefault(cn, {})
            dd[did] = dd.get(did, 0.0) + v
    return out


def counters_table(dbs: List[str], match: str = "") -> str:
    vals = counter_values(dbs)
    rx = re.compile(match) if match else None
    lines = []
    for kn in sorted(vals):
        if rx and not rx.search(kn):
            continue
        lines.append(kn)
        for cn in sorted(vals[kn]):
            d = vals[kn][cn]
            tot = sum(d.values())
            lines.append(f"    {cn:32s} total {tot:.4e}  per dispatch {tot / len(d):.4e}  "
                         f"({len(d)} dispatches)")
    return "\n".join(lines)


def _readings(vals, kernel_rx: str, counter: str):
    rx = re.compile(kernel_rx)
    per = [v for kn, cs in vals.items() if rx.search(kn) for v in cs.get(counter, {}).values()]
    if not per:
        raise SystemExit(f"no {counter} readings for kernels matching {kernel_rx!r}")
    return per


def _avg(vals, kernel_rx: str, counter: str):
    per = _readings(vals, kernel_rx, counter)
    return sum(per) / len(per), len(per)


def traffic(fetch_db: str, write_db: str, kernel: str, copy_rx: str, copy_bytes: float) -> str:
    fv, wv = counter_values([fetch_db]), counter_values([write_db])
    kx = "^" + re.escape(kernel) + r"(\b|$|<)"
    f_k, nf = _avg(fv, kx, "FETCH_SIZE")
    w_k, nw = _avg(wv, kx, "WRITE_SIZE")
    # the streaming dispatches read and write copy_bytes; other matches (a fill
    # of the source buffer) read nothing, so the largest reading is the one
    f_c = max(_readings(fv, copy_rx, "FETCH_SIZE"))
    w_c = max(_readings(wv, copy_rx, "WRITE_SIZE"))
    sf = copy_bytes / (f_c * 1024.0)
    sw = copy_bytes / (w_c * 1024.0)
    fetch_b, write_b = f_k * 1024.0 * sf, w_k * 1024.0 * sw
    return "\n".join([
        f"kernel: {kernel}",
        f"dispatches: fetch {nf}, write {nw}",
        f"FETCH_SIZE_KiB_per_dispatch: {f_k:.6e}",
        f"WRITE_SIZE_KiB_per_dispatch: {w_k:.6e}",
        f"copy_FETCH_SIZE_KiB: {f_c:.6e}",
        f"copy_WRITE_SIZE_KiB: {w_c:.6e}",
        f"copy_bytes_each_way: {copy_bytes:.6e}",
        f"fetch_scale: {sf:.4f}",
        f"write_scale: {sw:.4f}",
        f"fetch_bytes_per_launch: {fetch_b:.6e}",
        f"write_bytes_per_launch: {write_b:.6e}",
        f"bytes_per_launch: {fetch_b + write_b:.6e}",
    ])


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m nufht_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("trace")
    a.add_argument("db")
    a.add_argument("--match", default="")
    b = sub.add_parser("counters")
    b.add_argument("dbs", nargs="+")
    b.add_argument("--match", default="")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_db")
    t.add_argument("write_db")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    args = ap.parse_args(argv)
    if args.cmd == "trace":
        print(trace_table(args.db, args.match))
    elif args.cmd == "counters":
        print(counters_table(args.dbs, args.match))
    else:
        print(traffic(args.fetch_db, args.write_db, args.kernel, args.copy, args.copy_bytes))


if __name__ == "__main__":
    sys.exit(main())
