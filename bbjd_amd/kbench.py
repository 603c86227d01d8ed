"""Time one engine kernel in isolation (HIP events on a private stream).

    python -m bbjd_amd.kbench KERNEL N K [--reps R]

KERNEL: proj (Q^H x), proj2 (Q^H [x y]), update (x -= Q h), rotate (in-place
basis rotation k -> k/2, 8 N K (K/2) flops; --rotate-mode picks the kernel), cgs2 (one CGS2 with the second pass skipped on the
device), copy (plain copy of N complex values, the reference for
counter-measured HBM traffic), projout (the two-pass projection v -= Q Q^H v,
both passes always: for k <= 128 the fused form, first update and second
projection in one pass over Q), gram (G = Q^H Q, k x k, on the FP64 matrix
cores: also prints TFLOP/s for its 8 N K^2 flops).  Prints one line: kernel, shape, average us per
launch, algorithmic bytes per launch and GB/s.

qz: the projected step on the device, N problems of size K in one
bbjd_qz_device_batch launch (projected pencils of a random complex pencil,
nearest target, nkeep K/3), against the engine's host step for the same
problems on one thread (bbjd_qz_target: eigenvalues and target, then
bbjd_ritz_vector); prints us per launch, device problems/s and host ms per
problem.

gproj: the inner solve's complex64 projection Q^H x (k_proj plus its
finish) of --jobs independent solves (each its own Q, x and scratch),
recorded and flushed as grouped launches (launch.h) when --jobs > 1;
prints us per rep (host recording included), bytes per rep and GB/s.  Its
kernel trace gives the grouped kernel's own duration.

bcr: the exact preconditioner's solve (block cyclic reduction with the dense
rows' Schur border) on the serving workload's operator shape (block size 8,
block-tridiagonal, 4 dense rows, sigma 0.5+0.5i), N rows, K = 1 or 2
right-hand sides (bbjd_op_precond, bbjd_op_precond2), back to back on the
null stream; prints the average us per solve.  Its per-kernel figures come
from a kernel trace or counters around it.
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


def run_grouped(n, k, jobs, reps=50):
    """average ms per rep and bytes per rep of the grouped projection"""
    lib = _lib.load()
    lib.bbjd_kbench_grouped.restype = C.c_int
    lib.bbjd_kbench_grouped.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
    ms, by = C.c_double(), C.c_double()
    rc = lib.bbjd_kbench_grouped(int(jobs), int(n), int(k), int(reps), C.byref(ms), C.byref(by))
    if rc:
        raise RuntimeError(f"bbjd_kbench_grouped({jobs}, {n}, {k}) failed: {rc}")
    return ms.value, by.value


def run_bcr(n, nrhs, reps=50):
    """average ms per exact solve of nrhs right-hand sides"""
    import time

    import numpy as np

    from .pencil import BandedTauPencil, DeviceOperator
    if nrhs not in (1, 2):
        raise ValueError("bcr takes K = 1 or 2 right-hand sides")
    lib = _lib.load()
    op = DeviceOperator(BandedTauPencil.generate(n, 8, 1, 4, seed=3), 0.5 + 0.5j, pc="exact")
    try:
        if nrhs == 1:
            return op.time_apply("bbjd_op_precond", reps)
        rng = np.random.default_rng(0)
        xs = [_lib.DeviceBuffer.from_array(rng.standard_normal(n) + 1j * rng.standard_normal(n)) for _ in range(2)]
        ys = [_lib.DeviceBuffer(16 * n) for _ in range(2)]

        def solve():
            if lib.bbjd_op_precond2(n, xs[0].ptr, ys[0].ptr, xs[1].ptr, ys[1].ptr, op.handle):
                raise RuntimeError("bbjd_op_precond2 failed")
        solve()
        lib.bbjd_dev_sync()
        t = time.perf_counter()
        for _ in range(reps):
            solve()
        lib.bbjd_dev_sync()
        return (time.perf_counter() - t) * 1e3 / reps
    finally:
        op.free()


def _projected(k, seed, sigma):
    """W^H A V, W^H B V as the engine forms them (C = nu0 MA + mu0 MB upper
    triangular) for a random complex A, diagonal B and orthonormal V"""
    import numpy as np
    rng = np.random.default_rng(seed)
    n = 3 * k + 4
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = np.diag(1.0 + 0.1 * rng.standard_normal(n))
    V, _ = np.linalg.qr(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))
    nu0 = 1 / np.sqrt(1 + abs(sigma) ** 2)
    W, _ = np.linalg.qr((nu0 * A - sigma * nu0 * B) @ V)
    return W.conj().T @ A @ V, W.conj().T @ B @ V


def run_qz(nprob, k, reps=20):
    """(device us per launch of nprob problems, host ms per problem)"""
    import time

    import numpy as np

    from .qz_device import Batch
    lib = _lib.load()
    sigma = 0.3 - 0.2j
    pens = [_projected(k, 100 + i % 16, sigma) for i in range(nprob)]
    b = Batch([dict(MA=a, MB=m, sigma=sigma, nkeep=max(1, k // 3)) for a, m in pens])
    try:
        b.launch()
        b.results()
        t = time.perf_counter()
        for _ in range(reps):
            b.launch()
        lib.bbjd_dev_sync()
        dev_us = (time.perf_counter() - t) * 1e6 / reps
        st = b.results()[0]["status"]
    finally:
        b.free()
    f = lib.bbjd_qz_target
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, _lib.zcomplex, _lib.zcomplex, C.c_void_p,
                  C.c_void_p, _lib.zcomplex, C.c_int, C.c_void_p, C.c_int64, C.c_double, C.POINTER(C.c_int64)]
    lib.bbjd_ritz_vector.restype = C.c_int
    lib.bbjd_ritz_vector.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, _lib.zcomplex, _lib.zcomplex,
                                     C.c_void_p]
    lib.bbjd_qz_init.argtypes = [C.c_void_p, C.c_int64]
    lib.bbjd_qz_free.argtypes = [C.c_void_p]
    nu0 = 1 / np.sqrt(1 + abs(sigma) ** 2)
    al, be, x = (np.zeros(k, np.complex128) for _ in range(3))
    cv = np.zeros(1, np.complex128)
    p = C.c_int64(-1)
    hosts = min(nprob, 16)
    mats = [(np.asfortranarray(a), np.asfortranarray(m)) for a, m in pens[:hosts]]
    q = C.create_string_buffer(4096)
    # a fresh step state per problem: the state caches its last eigenvalues
    # while the pencil is unchanged
    t = time.perf_counter()
    for _ in range(reps):
        for MA, MB in mats:
            lib.bbjd_qz_init(q, k)
            rc = f(q, k, MA.ctypes.data, MB.ctypes.data, _lib.zc(nu0), _lib.zc(-sigma * nu0), al.ctypes.data,
                   be.ctypes.data, _lib.zc(sigma), 0, cv.ctypes.data, 0, 1e-10, C.byref(p))
            if rc == 0:
                lib.bbjd_ritz_vector(q, k, MA.ctypes.data, MB.ctypes.data, _lib.zc(al[p.value]),
                                     _lib.zc(be[p.value]), x.ctypes.data)
            lib.bbjd_qz_free(q)
    host_ms = (time.perf_counter() - t) * 1e3 / (reps * hosts)
    return dev_us, host_ms, st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("kernel", choices=sorted(KERNELS) + ["bcr", "qz", "gproj"])
    ap.add_argument("n", type=int)
    ap.add_argument("k", type=int)
    ap.add_argument("--reps", type=int, default=50)
    ap.add_argument("--proj-blocks", type=int, default=0, help="blocks per projection launch, 0 = default")
    ap.add_argument("--upd-rows", type=int, default=0, help="rows per block, 0 = chosen from the shape")
    ap.add_argument("--rotate-mode", type=int, default=1,
                    help="rotate: 1 streaming matrix-core form, 2 staged matrix-core form, 0 vector FP64")
    ap.add_argument("--rotate-chunk", type=int, default=0, help="rotate: input columns per LDS chunk (8, 16, 32)")
    ap.add_argument("--jobs", type=int, default=16, help="gproj: independent solves per grouped launch (1..16)")
    a = ap.parse_args()
    _lib.load().bbjd_set_rotate_mfma(a.rotate_mode)
    _lib.load().bbjd_set_rotate_chunk(a.rotate_chunk)
    if a.kernel == "qz":
        dev_us, host_ms, st = run_qz(a.n, a.k, a.reps)
        print(f"qz problems={a.n} k={a.k}: device {dev_us:.1f} us/launch ({a.n / (dev_us * 1e-6):.0f} problems/s, "
              f"status {st}); host step {host_ms:.3f} ms/problem on one thread "
              f"({1e3 / host_ms:.0f} problems/s)", flush=True)
        return
    if a.kernel == "gproj":
        ms, by = run_grouped(a.n, a.k, a.jobs, a.reps)
        print(f"gproj n={a.n} k={a.k} jobs={a.jobs}: {ms * 1e3:.2f} us/rep, {by / 1e6:.3f} MB/rep, "
              f"{by / (ms * 1e-3) / 1e9:.1f} GB/s", flush=True)
        return
    if a.kernel == "bcr":
        ms = run_bcr(a.n, a.k, a.reps)
        print(f"bcr n={a.n} rhs={a.k}: {ms * 1e3:.2f} us/solve", flush=True)
        return
    ms, by = run(a.kernel, a.n, a.k, a.reps, a.proj_blocks, a.upd_rows)
    extra = ""
    if a.kernel == "gram":
        extra = f", {8.0 * a.n * a.k * a.k / (ms * 1e-3) / 1e12:.2f} TFLOP/s"
    elif a.kernel == "rotate":
        keep = max(a.k // 2, 1)
        extra = f", {8.0 * a.n * a.k * keep / (ms * 1e-3) / 1e12:.2f} TFLOP/s"
    print(f"{a.kernel} n={a.n} k={a.k}: {ms * 1e3:.2f} us/launch, {by / 1e6:.3f} MB/launch, "
          f"{by / (ms * 1e-3) / 1e9:.1f} GB/s{extra}", flush=True)


if __name__ == "__main__":
    main()
