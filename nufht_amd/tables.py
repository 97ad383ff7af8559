"""Planning tables in the reference's cache format.

    python -m nufht_amd.tables write PATH
        computes every entry the reference's table generator fills
        (src/bounds.c:306-324) and writes them as its cache file does
        (src/fast_hankel_nufht.c:193-247); NUFHT_TABLE_CACHE may name the file

    python -m nufht_amd.tables show PATH NU TOL
        prints z_split and K_loc for K = 1..10 asymptotic terms from a cache
"""
from __future__ import annotations

import argparse
import math
import sys
from typing import Dict

import numpy as np

HEADER = (0x4E5546, 1, 401, 401, 32, 16)
TOLS = [10.0 ** -e for e in range(4, 16)]


def write(path: str) -> None:
    from . import _check, lib
    _check(lib().nufht_amd_write_table_cache(path.encode()), "nufht_amd_write_table_cache")


def load(path: str) -> Dict[str, np.ndarray]:
    """header, asy_coef[401, 64], asy_z[401, 32, 16], wimp_k[401, 32, 16]."""
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw, dtype=np.uint32, count=6)
    if tuple(int(v) for v in hdr) != HEADER:
        raise ValueError(f"{path}: not a table cache (header {list(hdr)})")
    o = 24
    coef = np.frombuffer(raw, dtype=np.float64, count=401 * 64, offset=o).reshape(401, 64)
    o += 401 * 64 * 8
    z = np.frombuffer(raw, dtype=np.float64, count=401 * 32 * 16, offset=o).reshape(401, 32, 16)
    o += 401 * 32 * 16 * 8
    k = np.frombuffer(raw, dtype=np.int32, count=401 * 32 * 16, offset=o).reshape(401, 32, 16)
    if o + 401 * 32 * 16 * 4 != len(raw):
        raise ValueError(f"{path}: {len(raw)} bytes, expected {o + 401 * 32 * 16 * 4}")
    return {"header": hdr, "asy_coef": coef, "asy_z": z, "wimp_k": k}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m nufht_amd.tables")
    sub = ap.add_subparsers(dest="cmd", required=True)
    w = sub.add_parser("write")
    w.add_argument("path")
    s = sub.add_parser("show")
    s.add_argument("path")
    s.add_argument("nu", type=int)
    s.add_argument("tol", type=float)
    args = ap.parse_args(argv)
    if args.cmd == "write":
        write(args.path)
        print("wrote", args.path)
        return 0
    t = load(args.path)
    j = max(1, math.ceil(-math.log10(args.tol) - 3)) - 1
    for K in range(1, 11):
        print(f"K_asy={K:2d}  z_split={t['asy_z'][args.nu, j, K - 1]:.10g}  "
              f"K_loc={t['wimp_k'][args.nu, j, K - 1]}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
