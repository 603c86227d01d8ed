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

Given a matrix to factor, rows with far more nonzeros than the rest (the
dense rows of bordered pencils) are pinned out of the sparse LU, as FPLUS
does (reference README.md:39-44): SuperLU on the whole of a banded pencil
with 4 dense rows fills U densely (32 M nonzeros at n = 8000, against
0.12 M for the band part).  The rest is factored alone and the pinned rows
enter through their Schur complement (include/bbjd_sparse.h,
bbjd_sparse_border).
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


class SparseBorder(C.Structure):
    _fields_ = [("d", C.c_int64), ("rows", C.c_void_p), ("M21", C.c_void_p), ("Y", C.c_void_p),
                ("Sinv", C.c_void_p)]


def dense_rows(M, max_rows=64, ratio=8.0, floor=32):
    """indices of the rows of M with far more nonzeros than the rest (more
    than max(floor, ratio x the median row)), when there are 1..max_rows of
    them; else an empty array"""
    M = sp.csr_matrix(M)
    nnz = np.diff(M.indptr)
    if nnz.size < 2:
        return np.zeros(0, np.int32)
    lim = max(floor, ratio * float(np.median(nnz)))
    rows = np.flatnonzero(nnz > lim).astype(np.int32)
    return rows if 1 <= rows.size <= max_rows else np.zeros(0, np.int32)


def pinned_factors(M, rows):
    """the bordered factorization that keeps dense rows out of the sparse
    LU: with R = rows and R' the rest, splu of M11 = M[R', R'], Y =
    M11^-1 M[R', R], M21 = M[R, R'] and S^-1 = (M[R, R] - M21 Y)^-1.  Raises
    RuntimeError when M11 is singular."""
    from scipy.sparse.linalg import splu
    M = sp.csr_matrix(M, dtype=np.complex128)
    n = M.shape[0]
    rows = np.asarray(rows, np.int32)
    rest = np.setdiff1d(np.arange(n, dtype=np.int32), rows)
    Mr = M[rest]
    lu11 = splu(sp.csc_matrix(Mr[:, rest]))
    Y = np.asfortranarray(lu11.solve(np.asarray(Mr[:, rows].todense(), dtype=np.complex128)))
    M21 = np.ascontiguousarray(M[rows][:, rest].toarray())
    S = np.asarray(M[rows][:, rows].todense(), dtype=np.complex128) - M21 @ Y
    return dict(rows=rows, rest=rest, lu11=lu11, Y=Y, M21=M21, Sinv=np.asfortranarray(np.linalg.inv(S)))


def bordered_solve(f, b):
    """M^-1 b on the host from pinned_factors' output (the device apply's
    arithmetic, for tests)"""
    b = np.asarray(b, np.complex128)
    z = f["lu11"].solve(b[f["rest"]])
    y2 = f["Sinv"] @ (b[f["rows"]] - f["M21"] @ z)
    y = np.empty_like(b)
    y[f["rest"]] = z - f["Y"] @ y2
    y[f["rows"]] = y2
    return y


def _setup(lib):
    if getattr(lib, "_sparse_ready", False):
        return
    lib.bbjd_sparse_op_create.restype = C.c_void_p
    lib.bbjd_sparse_op_create.argtypes = [C.POINTER(Csr), C.POINTER(Csr), C.POINTER(SparseLU), C.c_void_p]
    lib.bbjd_sparse_op_create_bordered.restype = C.c_void_p
    lib.bbjd_sparse_op_create_bordered.argtypes = [C.POINTER(Csr), C.POINTER(Csr), C.POINTER(SparseLU),
                                                   C.POINTER(SparseBorder), C.c_void_p]
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

    def __init__(self, A, B=None, preconditioner=None, stream=None, pin_dense_rows=True):
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
        lu = border = None
        self.pinned_rows = np.zeros(0, np.int32)
        if preconditioner is not None:
            f = preconditioner
            if sp.issparse(f) or isinstance(f, np.ndarray):
                # a matrix to factor: its dense rows (if any) are pinned out of
                # the sparse LU, which they would otherwise fill densely
                Mp = sp.csr_matrix(f, dtype=np.complex128)
                rows = dense_rows(Mp) if pin_dense_rows else np.zeros(0, np.int32)
                f = None
                if rows.size:
                    try:
                        bf = pinned_factors(Mp, rows)
                    except RuntimeError:
                        bf = None
                    if bf is not None:
                        f = bf["lu11"]
                        keep.append(bf)
                        border = SparseBorder(rows.size, bf["rows"].ctypes.data, bf["M21"].ctypes.data,
                                              bf["Y"].ctypes.data, bf["Sinv"].ctypes.data)
                        self.pinned_rows = bf["rows"]
                if f is None:
                    from scipy.sparse.linalg import splu
                    f = splu(sp.csc_matrix(Mp))
            if not all(hasattr(f, k) for k in ("L", "U", "perm_r", "perm_c")):
                raise ValueError("preconditioner must be a scipy SuperLU (splu) object or a sparse matrix")
            m = f.L.shape[0]
            pr = np.ascontiguousarray(f.perm_r, dtype=np.int32)
            pc = np.ascontiguousarray(f.perm_c, dtype=np.int32)
            keep += [pr, pc]
            lu = SparseLU(_csr(f.L, m, keep), _csr(f.U, m, keep), pr.ctypes.data, pc.ctypes.data)
        if border is not None:
            h = lib.bbjd_sparse_op_create_bordered(C.byref(a), C.byref(b) if b is not None else None, C.byref(lu),
                                                   C.byref(border), stream)
        else:
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
