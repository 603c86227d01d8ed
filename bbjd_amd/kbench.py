"""Time one engine kernel in isolation (HIP events on a private stream).

    python -m bbjd_amd.kbench KERNEL N K [--reps R]

KERNEL: proj (Q^H x), proj2 (Q^H [x y]), update (x -= Q h), rotate (in-place
basis rotation k -> k/2), cgs2 (one CGS2 with the second pass skipped on the
device), copy (plain copy of N complex values, the reference for
counter-measured HBM traffic), projout (the two-pass projection v -= Q Q^H v,
both passes always: for k <= 128 the fused form, first update and second
projection in one pass over Q), gram (G = Q^H Q, k x k, on the FP64 matrix
cores: also prints TFLOP/s for its 8 N K^2 flops).  Prints one line: kernel, shape, average us per
launch, algorithmic bytes per launch and GB/s.
"""
from __future__ import annotations

import argparse
import ctypes as C

from . import _lib

KERNELS = {"proj": 0, "proj2": 1, "update": 2, "rotate": 3, "cgs2": 4, "copy": 5, "projout": 6, "gram": 7}


def run(kernel, n, k, reps=50, proj_blocks=0, upd_rows=0, blocks_target=0):
    lib = _lib.load()
    if lib.bbjd_set_tiles(int(proj_blocks), int(upd_rows), int(blocks_target)):
        raise ValueError("update row tiles must be 0, 256, 512 or 1024")
    ms, by = C.c_double(), C.c_double()
    rc = lib.bbjd_kbench(KERNELS[kernel], int(n), int(k), int(reps), C.byref(ms), C.byref(by))
    if rc:
        raise RuntimeError(f"bbjd_kbench({kernel}, {n}, {k}) failed: {rc}")
    return ms.value, by.value


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kernel", choices=sorted(KERNELS))
    ap.add_argument("n", type=int)
    ap.add_argument("k", type=int)
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--proj-blocks", type=int, default=0, help="blocks per projection launch, 0 = default")
    ap.add_argument("--upd-rows", type=int, default=0, help="rows per block, 0 = chosen from the shape")
    a = ap.parse_args()
    ms, by = run(a.kernel, a.n, a.k, a.reps, a.proj_blocks, a.upd_rows)
    extra = ""
    if a.kernel == "gram":
        extra = f", {8.0 * a.n * a.k * a.k / (ms * 1e-3) / 1e12:.2f} TFLOP/s"
    print(f"{a.kernel} n={a.n} k={a.k}: {ms * 1e3:.2f} us/launch, {by / 1e6:.3f} MB/launch, "
          f"{by / (ms * 1e-3) / 1e9:.1f} GB/s{extra}", flush=True)


if __name__ == "__main__":
    main()
