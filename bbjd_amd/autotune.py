"""Tile autotuner for the engine's streaming kernels on the device at hand.

    python -m bbjd_amd.autotune [--out bbjd_amd/tiles_gfx950.json] [--reps 20]

For every (n, k) class of the table it times, in isolation (bbjd_kbench, HIP
events on a private stream), the projection Q^H x over candidate blocks per
launch and the update x -= Q h over candidate rows per block, keeps the
fastest of each, and writes the table as JSON.  ``apply(path)`` loads such a
file into the engine (bbjd_set_tile_table); ``bench.py`` applies the
committed table for this device when it exists.

Classes: n <= 3e4, 7e4, 1.5e5, 4e5, larger (timed at 2e4, 5e4, 1e5, 2e5,
1e6) and k <= 24, 48, 96, larger (timed at 16, 32, 64, 128): the shapes of
the serving workload and of the full-size configurations.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

from . import _lib

N_UPPER = [30000, 70000, 150000, 400000, 1 << 40]
N_AT = [20000, 50000, 100000, 200000, 1000000]
K_UPPER = [24, 48, 96, 4096]
K_AT = [16, 32, 64, 128]
PROJ_BLOCKS = [96, 128, 192, 256, 384, 512]
UPD_ROWS = [0, 64, 128, 256, 512, 1024]      # 0: the engine's shape rule
DEFAULT_FILE = os.path.join(os.path.dirname(__file__), "tiles_gfx950.json")


def _time(lib, which, n, k, reps):
    ms, by = C.c_double(), C.c_double()
    rc = lib.bbjd_kbench(which, int(n), int(k), int(reps), C.byref(ms), C.byref(by))
    if rc:
        raise RuntimeError(f"bbjd_kbench({which}, {n}, {k}) failed: {rc}")
    return ms.value * 1e3, by.value


def tune(reps=20, log=print):
    lib = _lib.load()
    lib.bbjd_set_tile_table(0, None, 0, None, None, None)
    proj = [[0] * len(K_AT) for _ in N_AT]
    upd = [[0] * len(K_AT) for _ in N_AT]
    measured = {}
    try:
        for i, n in enumerate(N_AT):
            for j, k in enumerate(K_AT):
                tp = {}
                for pb in PROJ_BLOCKS:
                    lib.bbjd_set_tiles(pb, 0, 0)
                    tp[pb], by = _time(lib, 0, n, k, reps)
                lib.bbjd_set_tiles(192, 0, 0)
                tu = {}
                for ur in UPD_ROWS:
                    lib.bbjd_set_tiles(0, ur, 0)
                    tu[ur], _ = _time(lib, 2, n, k, reps)
                lib.bbjd_set_tiles(0, 0, 0)
                proj[i][j] = min(tp, key=tp.get)
                upd[i][j] = min(tu, key=tu.get)
                measured[f"{n}x{k}"] = {"proj_us": tp, "update_us": tu}
                log(f"[autotune] n={n} k={k}: projection {proj[i][j]} blocks {tp[proj[i][j]]:.1f} us "
                    f"(192: {tp[192]:.1f}); update rows {upd[i][j]} {tu[upd[i][j]]:.1f} us (rule: {tu[0]:.1f})")
    finally:
        lib.bbjd_set_tiles(192, 0, 2048)
    return {"device": "gfx950", "tuned": time.strftime("%Y-%m-%d"), "reps": reps,
            "n_upper": N_UPPER, "k_upper": K_UPPER, "n_timed": N_AT, "k_timed": K_AT,
            "proj_blocks": proj, "upd_rows": upd, "measured_us": measured}


def apply(path=DEFAULT_FILE):
    """Load a tile table written by tune() into the engine; returns the table
    or None when the file does not exist."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    lib = _lib.load()
    nn, nk = len(t["n_upper"]), len(t["k_upper"])
    nu = (C.c_int64 * nn)(*t["n_upper"])
    ku = (C.c_int * nk)(*t["k_upper"])
    pb = (C.c_int * (nn * nk))(*[v for row in t["proj_blocks"] for v in row])
    ur = (C.c_int * (nn * nk))(*[v for row in t["upd_rows"] for v in row])
    if lib.bbjd_set_tile_table(nn, nu, nk, ku, pb, ur):
        raise ValueError(f"invalid tile table {path}")
    return t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_FILE)
    ap.add_argument("--reps", type=int, default=20)
    a = ap.parse_args()
    t = tune(a.reps, log=lambda s: print(s, file=sys.stderr, flush=True))
    with open(a.out, "w") as f:
        json.dump(t, f, indent=1)
    print(json.dumps({"written": a.out, "proj_blocks": t["proj_blocks"], "upd_rows": t["upd_rows"]}), flush=True)


if __name__ == "__main__":
    main()
