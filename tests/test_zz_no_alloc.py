"""The allocation-free CGS2 entry point (bbjd_block_orthogonalize_noalloc:
device arrays, reduction scratch in the library's file-scope device array)
against the allocating one (bbjd_block_orthogonalize on host arrays) on the
same inputs, through the C-ABI: two calls in a row, the second with a
smaller input, each with both DGKS branches.  Bar: the same kernels on the
same shape, so results agree to rounding (1e-13 of ||x||, the CGS2 parity
bar of the engine's tests is 1e-12)."""
import ctypes as C

import numpy as np
import pytest

from bbjd_amd import _lib

pytestmark = pytest.mark.gpu


def _orthonormal(n, k, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    q, _ = np.linalg.qr(M)
    return np.asfortranarray(q)


def _noalloc(lib, n, k, Q, x):
    dQ = _lib.DeviceBuffer.from_array(np.asfortranarray(Q).ravel(order="F")) if k else None
    dv = _lib.DeviceBuffer.from_array(x)
    dh = _lib.DeviceBuffer(max(k, 1) * 16)
    dn = _lib.DeviceBuffer(8)
    rc = lib.bbjd_block_orthogonalize_noalloc(n, k, dQ.ptr if dQ else None, n, dv.ptr, dh.ptr, dn.ptr, None)
    assert rc == 0
    assert lib.bbjd_dev_sync() == 0
    return dv.to_array(np.complex128, n), dh.to_array(np.complex128, max(k, 1))[:k], float(
        dn.to_array(np.float64, 1)[0])


def test_noalloc_cgs2_matches_allocating_entry():
    lib = _lib.load()
    lib.bbjd_block_orthogonalize_noalloc.restype = C.c_int
    lib.bbjd_block_orthogonalize_noalloc.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                                    C.c_void_p, C.c_void_p, C.c_void_p]
    for n, k in ((70001, 40), (3001, 9)):            # second call: a smaller input
        rng = np.random.default_rng(n + k)
        Q = _orthonormal(n, k, n)
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        for case, x in (("generic", v), ("near_parallel", Q @ rng.standard_normal(k) + 1e-6 * v)):
            vr, hr = x.copy(), np.zeros(k, np.complex128)
            nr = lib.bbjd_block_orthogonalize(n, k, Q.ctypes.data, vr.ctypes.data, hr.ctypes.data)
            vg, hg, ng = _noalloc(lib, n, k, Q, x)
            scale = np.linalg.norm(x)
            print(f"[noalloc] n={n} k={k} {case}: |dv| {np.linalg.norm(vg - vr) / scale:.1e} "
                  f"|dh| {np.linalg.norm(hg - hr) / scale:.1e} |dnorm| {abs(ng - nr) / scale:.1e}", flush=True)
            assert np.linalg.norm(vg - vr) <= 1e-13 * scale, (n, k, case)
            assert np.linalg.norm(hg - hr) <= 1e-13 * scale, (n, k, case)
            assert abs(ng - nr) <= 1e-13 * scale, (n, k, case)
            assert np.linalg.norm(Q.conj().T @ vg) <= 1e-12 * scale
    # sizes beyond the stated maximum are refused before anything is enqueued
    dv = _lib.DeviceBuffer(16)
    assert lib.bbjd_block_orthogonalize_noalloc((1 << 23) + 1, 1, dv.ptr, (1 << 23) + 1, dv.ptr, None, None,
                                                None) == -1
    assert lib.bbjd_block_orthogonalize_noalloc(16, 1025, dv.ptr, 16, dv.ptr, None, None, None) == -1
    assert lib.bbjd_block_orthogonalize_noalloc(16, 1, None, 16, dv.ptr, None, None, None) == -1
