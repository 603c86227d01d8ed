"""ORACLE access from Python -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liborc.so (the CPU restatement of the reference solver,
orc_jdqz.c, plus CPU operators for banded+tau pencils, orc_pencil.c) and,
when it was built, oracle/_ref/libbbjd_ref.so (the reference solver compiled
from its own sources).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_PATH = os.path.join(HERE, "liborc.so")
REF_PATH = os.path.join(HERE, "_ref", "libbbjd_ref.so")


class zcomplex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


CALLBACK = C.CFUNCTYPE(C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p)


class Params(C.Structure):
    """bbjd_params as laid out by reference include/bbjd.h:114-150."""
    _fields_ = [("n", C.c_int64), ("nev", C.c_int64), ("sigma", zcomplex), ("target", C.c_int),
                ("maxsub", C.c_int64), ("minsub", C.c_int64), ("tol", C.c_double),
                ("maxiter", C.c_int64), ("inner_maxiter", C.c_int64), ("inner_tol", C.c_double),
                ("v0", C.c_void_p), ("matvec_A", C.c_void_p), ("matvec_B", C.c_void_p),
                ("precond", C.c_void_p), ("op_data", C.c_void_p), ("verbose", C.c_int)]


class Result(C.Structure):
    _fields_ = [("nconv", C.c_int64), ("eigenvalues", C.c_void_p), ("eigenvectors", C.c_void_p),
                ("residuals", C.c_void_p), ("iters", C.c_int64), ("status", C.c_int)]


class TraceRow(C.Structure):
    _fields_ = [("iter", C.c_int64), ("kdim", C.c_int64), ("locked", C.c_int64),
                ("lam_re", C.c_double), ("lam_im", C.c_double), ("rel_res", C.c_double)]


_orc = None
_ref = None


def load_oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_PATH):
            raise ImportError(f"{ORC_PATH} not built (make -C oracle)")
        lib = C.CDLL(ORC_PATH)
        V, I64, D = C.c_void_p, C.c_int64, C.c_double
        lib.orc_solve.restype = C.c_int
        lib.orc_solve.argtypes = [C.POINTER(Params), C.POINTER(Result), C.POINTER(TraceRow), I64,
                                  C.POINTER(I64), C.POINTER(I64)]
        lib.orc_result_free.argtypes = [C.POINTER(Result)]
        lib.orc_block_orthogonalize.restype = D
        lib.orc_block_orthogonalize.argtypes = [I64, I64, V, V, V]
        lib.orc_project_out.argtypes = [I64, I64, V, V]
        lib.orc_zlartg.argtypes = [zcomplex, zcomplex, C.POINTER(D), C.POINTER(zcomplex),
                                   C.POINTER(zcomplex)]
        lib.orc_pencil_create.restype = V
        lib.orc_pencil_create.argtypes = [I64, C.c_int, C.c_int, C.c_int, V, V, V, zcomplex, C.c_int]
        lib.orc_pencil_free.argtypes = [V]
        lib.orc_pencil_dense.argtypes = [V, V, V]
        for nm in ("orc_matvec_A", "orc_matvec_B", "orc_precond"):
            getattr(lib, nm).restype = C.c_int
            getattr(lib, nm).argtypes = [I64, V, V, V]
        lib.orc_set_threads.argtypes = [C.c_int]
        _orc = lib
    return _orc


def have_ref():
    return os.path.exists(REF_PATH)


def load_ref():
    global _ref
    if _ref is None:
        if not have_ref():
            return None
        lib = C.CDLL(REF_PATH)
        lib.bbjd_params_init.argtypes = [C.POINTER(Params)]
        lib.bbjd_workspace_alloc.restype = C.c_void_p
        lib.bbjd_workspace_alloc.argtypes = [C.POINTER(Params)]
        lib.bbjd_workspace_free.argtypes = [C.c_void_p]
        lib.bbjd_solve.restype = C.c_int
        lib.bbjd_solve.argtypes = [C.POINTER(Params), C.c_void_p, C.POINTER(Result)]
        lib.bbjd_result_free.argtypes = [C.POINTER(Result)]
        lib.bbjd_block_orthogonalize.restype = C.c_double
        lib.bbjd_block_orthogonalize.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                                                 C.c_void_p]
        _ref = lib
    return _ref


def _view(addr, n, writable):
    a = np.frombuffer((C.c_double * (2 * n)).from_address(addr), dtype=np.complex128)
    if not writable:
        a.flags.writeable = False
    return a


def py_callback(fn):
    """ctypes callback around a Python y = fn(x)."""
    def cb(n, x, y, data):
        try:
            _view(y, n, True)[:] = np.asarray(fn(_view(x, n, False)), dtype=np.complex128)
            return 0
        except Exception:  # noqa: BLE001
            return -1
    return CALLBACK(cb)


class CpuPencilOps:
    """orc_pencil: CPU matvecs and preconditioner of a banded+tau pencil
    (pc 0 none, 1 exact LU via zgbtrf + Schur complement, 2 block Jacobi)."""

    def __init__(self, pencil, sigma, pc=1):
        lib = load_oracle()
        self.pencil = pencil          # keeps the arrays alive
        self.handle = lib.orc_pencil_create(
            pencil.n, pencil.bs, pencil.hw, pencil.ntau, pencil.band.ctypes.data,
            pencil.tau.ctypes.data if pencil.ntau else None, pencil.bdiag.ctypes.data,
            zcomplex(complex(sigma).real, complex(sigma).imag), int(pc))
        if not self.handle:
            raise RuntimeError("orc_pencil_create failed")
        self.pc = pc

    def addr(self, name):
        return C.cast(getattr(load_oracle(), name), C.c_void_p).value

    def apply(self, name, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty_like(x)
        rc = getattr(load_oracle(), name)(x.size, x.ctypes.data, y.ctypes.data, self.handle)
        if rc:
            raise RuntimeError(name)
        return y

    def __del__(self):
        try:
            load_oracle().orc_pencil_free(self.handle)
        except Exception:
            pass


def gmres(n, QR, QL, ta, tb, r, ops, maxiter, tol):
    """orc_correction_gmres on host arrays: QR, QL are n x ndefl."""
    lib = load_oracle()
    QR = np.asfortranarray(QR, dtype=np.complex128)
    QL = np.asfortranarray(QL, dtype=np.complex128)
    ndefl = QR.shape[1]
    r = np.ascontiguousarray(r, dtype=np.complex128)
    z = np.zeros(n, np.complex128)
    V = np.zeros((maxiter + 1) * n, np.complex128)
    H = np.zeros((maxiter + 1) * maxiter, np.complex128)
    cs = np.zeros(maxiter, np.float64)
    sn = np.zeros(maxiter, np.complex128)
    g = np.zeros(maxiter + 1, np.complex128)
    t1 = np.zeros(n, np.complex128)
    t2 = np.zeros(n, np.complex128)
    steps = C.c_int64(0)
    f = lib.orc_correction_gmres
    f.restype = C.c_int
    f.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, zcomplex, zcomplex, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_double,
                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.POINTER(C.c_int64)]
    rc = f(n, ndefl, QR.ctypes.data, QL.ctypes.data, zcomplex(ta.real, ta.imag),
           zcomplex(tb.real, tb.imag), r.ctypes.data, z.ctypes.data, ops.addr("orc_matvec_A"),
           ops.addr("orc_matvec_B"), ops.addr("orc_precond") if ops.pc else None, ops.handle,
           maxiter, tol, V.ctypes.data, H.ctypes.data, cs.ctypes.data, sn.ctypes.data,
           g.ctypes.data, t1.ctypes.data, t2.ctypes.data, C.byref(steps))
    if rc:
        raise RuntimeError("orc_correction_gmres")
    return z


def _fill_params(p, n, nev, sigma, target, tol, maxiter, maxsub, minsub, inner_maxiter,
                 inner_tol, verbose, v0):
    p.n, p.nev = n, nev
    p.sigma = zcomplex(complex(sigma).real, complex(sigma).imag)
    p.target, p.tol, p.maxiter = target, tol, maxiter
    p.maxsub, p.minsub, p.inner_maxiter, p.inner_tol, p.verbose = maxsub, minsub, inner_maxiter, \
        inner_tol, verbose
    p.v0 = v0.ctypes.data if v0 is not None else None


def _collect(res, n):
    nconv = int(res.nconv)
    ev = np.array(_view(res.eigenvalues, nconv, False)) if nconv else np.empty(0, np.complex128)
    vec = (np.array(_view(res.eigenvectors, n * nconv, False)).reshape(nconv, n).T
           if nconv else np.empty((n, 0), np.complex128))
    rs = (np.frombuffer((C.c_double * nconv).from_address(res.residuals), dtype=np.float64).copy()
          if nconv else np.empty(0))
    return ev, vec, rs, nconv, int(res.iters), int(res.status)


def solve(which, n, nev, sigma=0.0, target=0, tol=1e-10, maxiter=200, maxsub=0, minsub=0,
          inner_maxiter=0, inner_tol=1e-2, verbose=0, v0=None, ops=None, matvec_A=None,
          matvec_B=None, precond=None, trace_cap=0):
    """Run the oracle restatement ('oracle') or the compiled reference ('ref').

    Operators: ``ops`` (a CpuPencilOps: C callbacks) or Python callables.
    Returns dict(evals, evecs, resid, nconv, iters, status[, trace]).
    """
    keep = []
    if v0 is not None:
        v0 = np.ascontiguousarray(v0, dtype=np.complex128)
    if ops is not None:
        fA, fB = ops.addr("orc_matvec_A"), ops.addr("orc_matvec_B")
        fP = ops.addr("orc_precond") if ops.pc else None
        od = ops.handle
    else:
        cA, cB = py_callback(matvec_A), py_callback(matvec_B)
        keep += [cA, cB]
        fA, fB = C.cast(cA, C.c_void_p).value, C.cast(cB, C.c_void_p).value
        fP = None
        if precond is not None:
            cP = py_callback(precond)
            keep.append(cP)
            fP = C.cast(cP, C.c_void_p).value
        od = None
    p = Params()
    res = Result()
    out = {}
    if which == "oracle":
        lib = load_oracle()
        C.memset(C.byref(p), 0, C.sizeof(p))
        _fill_params(p, n, nev, sigma, target, tol, maxiter, maxsub, minsub, inner_maxiter,
                     inner_tol, verbose, v0)
        p.matvec_A, p.matvec_B, p.precond, p.op_data = fA, fB, fP, od
        tr = (TraceRow * max(trace_cap, 1))()
        ntr, inner = C.c_int64(0), C.c_int64(0)
        lib.orc_solve(C.byref(p), C.byref(res), tr, trace_cap, C.byref(ntr), C.byref(inner))
        ev, vec, rs, nconv, iters, status = _collect(res, n)
        lib.orc_result_free(C.byref(res))
        out["trace"] = np.array([(r.iter, r.kdim, r.locked, r.lam_re, r.lam_im, r.rel_res)
                                 for r in tr[: ntr.value]])
        out["inner_steps"] = inner.value
    elif which == "ref":
        lib = load_ref()
        if lib is None:
            raise ImportError("oracle/_ref not built")
        lib.bbjd_params_init(C.byref(p))
        _fill_params(p, n, nev, sigma, target, tol, maxiter, maxsub, minsub, inner_maxiter,
                     inner_tol, verbose, v0)
        p.matvec_A, p.matvec_B, p.precond, p.op_data = fA, fB, fP, od
        ws = lib.bbjd_workspace_alloc(C.byref(p))
        if not ws:
            raise MemoryError("reference workspace")
        lib.bbjd_solve(C.byref(p), ws, C.byref(res))
        lib.bbjd_workspace_free(ws)
        ev, vec, rs, nconv, iters, status = _collect(res, n)
        lib.bbjd_result_free(C.byref(res))
    else:
        raise ValueError(which)
    out.update(evals=ev, evecs=vec, resid=rs, nconv=nconv, iters=iters, status=status)
    return out
