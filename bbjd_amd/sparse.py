"""General sparse pencils on the GPU (include/bbjd_sparse.h).

``SparseDeviceOperator(A, B, preconditioner)`` puts A and B (scipy sparse
matrices of any sparsity; B None is the identity) in HBM as CSR and, with a
preconditioner, the factors of a sparse LU of it: a scipy ``SuperLU``
(``splu(A - sigma B)``, what the reference's own tests pass, reference
tests/test_bbjd.py:59-73) or a sparse matrix to factor with ``splu``.  The
factorization runs once on the host; every apply (matvecs and the two
triangular solves) runs on the GPU, so an iteration moves no n-vector over
PCIe.  Pass it to ``jdqz`` as A, or call ``jdqz(A, B, ...,
preconditioner=lu, device=True)``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib


class Csr(C.Structure):
    _fields_ = [("n", C.c_int64), ("ncols", C.c_int64), ("ptr", C.c_void_p), ("col", C.c_void_p),
                ("val", C.c_void_p)]


class SparseLU(C.Structure):
    _fields_ = [("L", Csr), ("U", Csr), ("perm_r", C.c_void_p), ("perm_c", C.c_void_p)]


def _setup(lib):
    if getattr(lib, "_sparse_ready", False):
        return
    lib.bbjd_sparse_op_create.restype = C.c_void_p
    lib.bbjd_sparse_op_create.argtypes = [C.POINTER(Csr), C.POINTER(Csr), C.POINTER(SparseLU), C.c_void_p]
    lib.bbjd_sparse_op_free.argtypes = [C.c_void_p]
    lib.bbjd_sparse_op_bytes.restype = C.c_size_t
    lib.bbjd_sparse_op_bytes.argtypes = [C.c_void_p]
    lib.bbjd_sparse_op_levels.restype = C.c_int
    lib.bbjd_sparse_op_levels.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64)]
    lib.bbjd_sparse_levels.restype = C.c_int64
    lib.bbjd_sparse_levels.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    for nm in ("bbjd_sparse_matvec_A", "bbjd_sparse_matvec_B", "bbjd_sparse_precond"):
        getattr(lib, nm).restype = C.c_int
        getattr(lib, nm).argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    lib._sparse_ready = True


def _csr(M, n, keep):
    """host CSR arrays of M (complex128 values, int64 offsets, int32
    columns) and the bbjd_csr describing them"""
    M = sp.csr_matrix(M, dtype=np.complex128)
    if M.shape != (n, n):
        raise ValueError(f"matrix must be {n} x {n}; got {M.shape}")
    ptr = np.ascontiguousarray(M.indptr, dtype=np.int64)
    col = np.ascontiguousarray(M.indices, dtype=np.int32)
    val = np.ascontiguousarray(M.data, dtype=np.complex128)
    keep += [ptr, col, val]
    return Csr(n, n, ptr.ctypes.data, col.ctypes.data, val.ctypes.data)


def levels(F, lower):
    """level of each row of a triangular sparse factor (the solve plan's
    host logic, bbjd_sparse_levels): -1 when F is not triangular"""
    lib = _lib.load()
    _setup(lib)
    F = sp.csr_matrix(F)
    n = F.shape[0]
    ptr = np.ascontiguousarray(F.indptr, dtype=np.int64)
    col = np.ascontiguousarray(F.indices, dtype=np.int32)
    out = np.zeros(n, np.int32)
    nlev = lib.bbjd_sparse_levels(n, ptr.ctypes.data, col.ctypes.data, int(bool(lower)), out.ctypes.data)
    return int(nlev), out


class SparseDeviceOperator:
    """A general sparse pencil (A, B) on the GPU with an optional sparse-LU
    preconditioner (see the module docstring)."""

    c64 = False      # no complex64 copies: a mixed inner solve converts around the FP64 applies

    def __init__(self, A, B=None, preconditioner=None, stream=None):
        lib = _lib.load()
        _setup(lib)
        if not sp.issparse(A):
            A = sp.csr_matrix(np.asarray(A))
        n = A.shape[0]
        if A.shape != (n, n):
            raise ValueError(f"A must be square; got shape {A.shape}")
        keep = []
        a = _csr(A, n, keep)
        b = None if B is None else _csr(B, n, keep)
        lu = None
        if preconditioner is not None:
            f = preconditioner
            if sp.issparse(f) or isinstance(f, np.ndarray):
                from scipy.sparse.linalg import splu
                f = splu(sp.csc_matrix(f, dtype=np.complex128))
            if not all(hasattr(f, k) for k in ("L", "U", "perm_r", "perm_c")):
                raise ValueError("preconditioner must be a scipy SuperLU (splu) object or a sparse matrix")
            pr = np.ascontiguousarray(f.perm_r, dtype=np.int32)
            pc = np.ascontiguousarray(f.perm_c, dtype=np.int32)
            keep += [pr, pc]
            lu = SparseLU(_csr(f.L, n, keep), _csr(f.U, n, keep), pr.ctypes.data, pc.ctypes.data)
        h = lib.bbjd_sparse_op_create(C.byref(a), C.byref(b) if b is not None else None,
                                      C.byref(lu) if lu is not None else None, stream)
        if not h:
            raise RuntimeError("bbjd_sparse_op_create failed (invalid factors or device memory)")
        self.handle = h
        self.n = n
        self.has_precond = lu is not None

    @property
    def shape(self):
        return (self.n, self.n)

    @property
    def device_bytes(self):
        return _lib.load().bbjd_sparse_op_bytes(self.handle)

    def plan(self):
        """(levels of L, levels of U, launches per preconditioner apply)"""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _lib.load().bbjd_sparse_op_levels(self.handle, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def device_callbacks(self, precision, with_precond):
        """the engine's callback slots this operator fills (C function names)"""
        cb = {"matvec_A": "bbjd_sparse_matvec_A", "matvec_B": "bbjd_sparse_matvec_B"}
        if with_precond:
            if not self.has_precond:
                raise ValueError("this SparseDeviceOperator was built without a preconditioner")
            cb["precond"] = "bbjd_sparse_precond"
        return cb

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

    def matvec(self, x):
        return self._apply("bbjd_sparse_matvec_A", x)

    def matvec_B(self, x):
        return self._apply("bbjd_sparse_matvec_B", x)

    def solve(self, x):
        return self._apply("bbjd_sparse_precond", x)

    def free(self):
        if getattr(self, "handle", None):
            _lib.load().bbjd_sparse_op_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
