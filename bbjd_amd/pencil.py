"""Block-banded pencils with dense tau rows, and their device operators.

``BandedTauPencil`` holds the host arrays in the layout of
include/bbjd_ops.h; ``generate`` draws the seeded benchmark pencils (the C
generator of the library, so Python, C and the benchmarks see identical
matrices).  ``DeviceOperator`` uploads a pencil and builds the device
matvec/preconditioner for a shift; its C callbacks are what the engine calls
in DEVICE callback space.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class BandedTauPencil:
    """A = block-banded (block size bs, block half-width hw) with the last
    ntau rows dense; B = diag(bdiag)."""

    def __init__(self, n, bs, hw, ntau, band, tau, bdiag):
        self.n, self.bs, self.hw, self.ntau = int(n), int(bs), int(hw), int(ntau)
        if self.n % self.bs:
            raise ValueError("n must be a multiple of the block size")
        nb, nd = self.n // self.bs, 2 * self.hw + 1
        self.band = np.ascontiguousarray(band, dtype=np.complex128).reshape(nb * nd * bs * bs)
        self.tau = np.ascontiguousarray(tau, dtype=np.complex128).reshape(max(self.ntau, 0) * self.n)
        self.bdiag = np.ascontiguousarray(bdiag, dtype=np.complex128).reshape(self.n)

    @classmethod
    def generate(cls, n, bs, hw, ntau, seed, diag_shift=4.0):
        lib = _lib.load()
        n, bs, hw, ntau = int(n), int(bs), int(hw), int(ntau)
        nb, nd = n // bs, 2 * hw + 1
        band = np.empty(nb * nd * bs * bs, dtype=np.complex128)
        tau = np.empty(max(ntau, 1) * n, dtype=np.complex128)
        bdiag = np.empty(n, dtype=np.complex128)
        rc = lib.bbjd_pencil_generate(n, bs, hw, ntau, C.c_uint64(seed), float(diag_shift),
                                      band.ctypes.data, tau.ctypes.data, bdiag.ctypes.data)
        if rc:
            raise ValueError("invalid pencil shape")
        return cls(n, bs, hw, ntau, band, tau[: ntau * n], bdiag)

    @property
    def n1(self):
        return self.n - self.ntau

    def desc(self):
        d = _lib.PencilDesc()
        d.n, d.bs, d.hw, d.ntau = self.n, self.bs, self.hw, self.ntau
        d.band = self.band.ctypes.data
        d.tau = self.tau.ctypes.data if self.ntau else None
        d.bdiag = self.bdiag.ctypes.data
        return d

    def to_scipy(self):
        """(A, B) as scipy CSR matrices (for host-side checks)."""
        import scipy.sparse as sp
        n, bs, hw = self.n, self.bs, self.hw
        nb, nd = n // bs, 2 * hw + 1
        blk = self.band.reshape(nb, nd, bs, bs)         # [bi, d, c, r]
        rows, cols, vals = [], [], []
        bi, d, c, r = np.meshgrid(np.arange(nb), np.arange(nd), np.arange(bs), np.arange(bs),
                                  indexing="ij")
        bj = bi + d - hw
        i = bi * bs + r
        j = bj * bs + c
        keep = (bj >= 0) & (bj < nb) & (i < self.n1)
        rows.append(i[keep]); cols.append(j[keep]); vals.append(blk[keep])
        if self.ntau:
            t = self.tau.reshape(self.ntau, n)
            ti, tj = np.meshgrid(np.arange(self.ntau), np.arange(n), indexing="ij")
            rows.append((self.n1 + ti).ravel()); cols.append(tj.ravel()); vals.append(t.ravel())
        A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(n, n))
        B = sp.diags(self.bdiag, format="csr")
        return A, B


_PC = {"none": _lib.PC_NONE, None: _lib.PC_NONE, "exact": _lib.PC_EXACT_LU,
       "exact_lu": _lib.PC_EXACT_LU, "block_jacobi": _lib.PC_BLOCK_JACOBI}


class DeviceOperator:
    """The pencil on the GPU with the preconditioner of A - sigma B.

    pc: 'exact' (block cyclic reduction + Schur complement of the tau rows),
    'block_jacobi' (bs x bs diagonal-block inverses) or 'none'.
    c64: also keep complex64 copies for the mixed-precision inner solve.
    """

    def __init__(self, pencil: BandedTauPencil, sigma=0.0, pc="exact", c64=False, stream=None):
        lib = _lib.load()
        if pc not in _PC:
            raise ValueError(f"pc must be one of {list(k for k in _PC if k)}")
        self.pencil = pencil
        self.n = pencil.n
        self.sigma = complex(sigma)
        self.pc = _PC[pc]
        self.c64 = bool(c64)
        self._desc = pencil.desc()
        h = lib.bbjd_op_create(C.byref(self._desc), _lib.zc(sigma), self.pc, int(self.c64), stream)
        if not h:
            raise RuntimeError("bbjd_op_create failed (device memory or a singular pivot block)")
        self.handle = h

    @property
    def shape(self):
        return (self.n, self.n)

    @property
    def device_bytes(self):
        return _lib.load().bbjd_op_bytes(self.handle)

    def device_callbacks(self, precision, with_precond):
        """the engine's callback slots this operator fills (C function names):
        the three callbacks, the fused A/B and residual passes, both
        preconditioner solves in one pass, and the complex64 forms for a
        mixed inner solve when the operator keeps complex64 copies"""
        cb = {"matvec_A": "bbjd_op_matvec_A", "matvec_B": "bbjd_op_matvec_B",
              "matvec_AB": "bbjd_op_matvec_AB", "matvec_ABr": "bbjd_op_matvec_ABr"}
        if with_precond:
            cb.update(precond="bbjd_op_precond", precond2="bbjd_op_precond2")
        if precision == "mixed" and self.c64:
            cb.update(matvec_A_c64="bbjd_op_matvec_A_c64", matvec_B_c64="bbjd_op_matvec_B_c64",
                      matvec_AB_c64="bbjd_op_matvec_AB_c64")
            if with_precond:
                cb["precond_c64"] = "bbjd_op_precond_c64"
        return cb

    def free(self):
        if getattr(self, "handle", None):
            _lib.load().bbjd_op_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # host-array convenience applies (copy in, apply on the device, copy out)
    def _apply(self, name, x):
        lib = _lib.load()
        x = np.ascontiguousarray(x, dtype=np.complex128)
        dx = _lib.DeviceBuffer.from_array(x)
        dy = _lib.DeviceBuffer(x.nbytes)
        rc = getattr(lib, name)(self.n, dx.ptr, dy.ptr, self.handle)
        rc |= lib.bbjd_dev_sync()
        if rc:
            raise RuntimeError(f"{name} failed")
        return dy.to_array(np.complex128, self.n)

    def time_apply(self, name, reps=20):
        """average ms of a device apply (HIP events, null stream)"""
        import ctypes as C
        lib = _lib.load()
        x = np.random.default_rng(0).standard_normal(self.n) + 0j
        dx = _lib.DeviceBuffer.from_array(x)
        dy = _lib.DeviceBuffer(x.nbytes)
        ms = C.c_double()
        rc = lib.bbjd_time_op(_lib.fn_addr(name), self.n, dx.ptr, dy.ptr, self.handle, reps, C.byref(ms))
        if rc:
            raise RuntimeError(f"timing {name} failed")
        return ms.value

    def matvec(self, x):
        return self._apply("bbjd_op_matvec_A", x)

    def matvec_B(self, x):
        return self._apply("bbjd_op_matvec_B", x)

    def solve(self, x):
        return self._apply("bbjd_op_precond", x)

    def solve_c64(self, x):
        """the complex64 preconditioner apply of the mixed inner solve
        (bbjd_op_precond_c64; the operator needs c64=True)"""
        lib = _lib.load()
        x = np.ascontiguousarray(x, dtype=np.complex64)
        dx = _lib.DeviceBuffer.from_array(x)
        dy = _lib.DeviceBuffer(x.nbytes)
        rc = lib.bbjd_op_precond_c64(self.n, dx.ptr, dy.ptr, self.handle)
        rc |= lib.bbjd_dev_sync()
        if rc:
            raise RuntimeError("bbjd_op_precond_c64 failed")
        return dy.to_array(np.complex64, self.n)

    def matvec_AB(self, x, single=False):
        """(A x, B x) in one pass over x (bbjd_op_matvec_AB; single: the
        complex64 form, bbjd_op_matvec_AB_c64)"""
        lib = _lib.load()
        dt = np.complex64 if single else np.complex128
        x = np.ascontiguousarray(x, dtype=dt)
        dx = _lib.DeviceBuffer.from_array(x)
        da, db = _lib.DeviceBuffer(x.nbytes), _lib.DeviceBuffer(x.nbytes)
        fn = lib.bbjd_op_matvec_AB_c64 if single else lib.bbjd_op_matvec_AB
        rc = fn(self.n, dx.ptr, da.ptr, db.ptr, self.handle)
        rc |= lib.bbjd_dev_sync()
        if rc:
            raise RuntimeError("bbjd_op_matvec_AB failed")
        return da.to_array(dt, self.n), db.to_array(dt, self.n)

    def matvec_ABr(self, x, alpha, beta):
        """(A x, B x, r = alpha A x + beta B x, ||r||) in one pass over x
        (bbjd_op_matvec_ABr)"""
        lib = _lib.load()
        x = np.ascontiguousarray(x, dtype=np.complex128)
        dx = _lib.DeviceBuffer.from_array(x)
        da, db, dr = _lib.DeviceBuffer(x.nbytes), _lib.DeviceBuffer(x.nbytes), _lib.DeviceBuffer(x.nbytes)
        dn = _lib.DeviceBuffer(8)
        rc = lib.bbjd_op_matvec_ABr(self.n, dx.ptr, da.ptr, db.ptr, _lib.zc(alpha), _lib.zc(beta), dr.ptr, dn.ptr,
                                    self.handle)
        rc |= lib.bbjd_dev_sync()
        if rc:
            raise RuntimeError("bbjd_op_matvec_ABr failed")
        return (da.to_array(np.complex128, self.n), db.to_array(np.complex128, self.n),
                dr.to_array(np.complex128, self.n), float(dn.to_array(np.float64, 1)[0]))

    def solve2(self, x0, x1):
        """both preconditioner solves of one correction step in one pass
        (bbjd_op_precond2)"""
        lib = _lib.load()
        x0 = np.ascontiguousarray(x0, dtype=np.complex128)
        x1 = np.ascontiguousarray(x1, dtype=np.complex128)
        d0, d1 = _lib.DeviceBuffer.from_array(x0), _lib.DeviceBuffer.from_array(x1)
        e0, e1 = _lib.DeviceBuffer(x0.nbytes), _lib.DeviceBuffer(x1.nbytes)
        rc = lib.bbjd_op_precond2(self.n, d0.ptr, e0.ptr, d1.ptr, e1.ptr, self.handle)
        rc |= lib.bbjd_dev_sync()
        if rc:
            raise RuntimeError("bbjd_op_precond2 failed")
        return e0.to_array(np.complex128, self.n), e1.to_array(np.complex128, self.n)
