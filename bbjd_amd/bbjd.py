"""``jdqz``: scipy-eigs-style Jacobi-Davidson QZ on one MI355X.

Mirrors the reference wrapper (kyle-augustson/BBJD bbjd.py:59-277): same
signature, defaults, input handling, error mapping and return values, over
the GPU engine (libbbjd.so).

    evals, evecs = jdqz(A, B, k=10, sigma=0.5, preconditioner=lu)

A and B may be dense arrays (-> CSR), scipy sparse matrices (CSC matvec),
objects with .matvec, or callables (reference bbjd.py:161-205); the
preconditioner an object with .solve or a callable (bbjd.py:208-219).  These
run on the host and the engine stages each vector to/from HBM.

Device fast path (extension): pass a ``DeviceOperator`` (banded + dense
rows) or a ``SparseDeviceOperator`` (any sparsity, sparse-LU
preconditioner) as A (B None) and the same operator, or True, as
``preconditioner``; or keep the reference's call with scipy matrices and a
SuperLU preconditioner and add ``device=True``.  Then the matvecs and the
preconditioner run on the GPU as well and no n-vector crosses PCIe during
the iteration.  ``precision='mixed'`` runs the inner
GMRES (inner_maxiter > 0) in complex64 with FP32 accumulation
(``inner_form='projected'``: on the projected preconditioner, one oblique
projection per step), and ``graphs=True`` replays its steps from captured
HIP graphs.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _bbjd_c
from .pencil import DeviceOperator
from .sparse import SparseDeviceOperator

# Target values fixed by reference include/bbjd.h:83-91 / bbjd.py:48-56.
_TARGET_MAP = {
    'nearest': 0,
    'LM': 1,
    'SM': 2,
    'LR': 3,
    'SR': 4,
    'LI': 5,
    'SI': 6,
}

_STATUS_MSG = {
    -1: "memory allocation failed",
    -2: "invalid parameters",
    -3: "callback error (matvec or preconditioner failed)",
    -4: "LAPACK error in projected eigenproblem",
    -5: "subspace breakdown",
}


def _identity(x):
    """B = I (B omitted): a copy, since the engine owns its output buffer"""
    return np.array(x, dtype=np.complex128)


def _from_sparse(M):
    Mc = sp.csc_matrix(M, dtype=np.complex128)
    return Mc.dot


def _from_method(name):
    def build(obj):
        method = getattr(obj, name)
        return lambda x: np.asarray(method(x), dtype=np.complex128)
    return build


# (test, builder) pairs, first match wins: what a matvec operand may be
# (reference bbjd.py:183-205) and what a preconditioner may be (:208-219)
_OPERAND_KINDS = (
    (sp.issparse, _from_sparse),
    (lambda o: hasattr(o, "matvec"), _from_method("matvec")),
    (callable, lambda f: f),
)
_PRECOND_KINDS = (
    (lambda o: hasattr(o, "solve"), _from_method("solve")),
    (callable, lambda f: f),
)


def _adapt(obj, kinds, error):
    for test, build in kinds:
        if test(obj):
            return build(obj)
    raise ValueError(error)


def _start_vector(v0, n):
    if v0 is None:
        return None
    v0 = np.ascontiguousarray(v0, dtype=np.complex128)
    if v0.shape != (n,):
        raise ValueError(f"v0 must have shape ({n},); got {v0.shape}")
    return v0


def jdqz(A, B=None, k: int = 6, sigma: complex = 0.0, which: str = 'nearest', tol: float = 1e-10,
         maxiter: int = 200, maxsub: int = 0, minsub: int = 0, inner_maxiter: int = 0,
         inner_tol: float = 1e-2, preconditioner=None, v0=None, verbose: int = 0,
         return_info: bool = False, precision: str = 'fp64', graphs: bool = False, device: bool = False,
         inner_form: str = 'reference'):
    """Jacobi-Davidson QZ for A x = lambda B x, eigenvalues nearest sigma
    (or by `which`).

    maxsub = 0 picks min(max(3k + 10, 40), n) and minsub = 0 picks
    max(2k + 5, 20) (the C defaults, reference src/jd_core.c:95-98, 750-752).
    Returns (eigenvalues, eigenvectors) or, with return_info, also
    {'nconv', 'iters', 'status', 'residuals'}.  Raises ValueError on invalid
    input and RuntimeError on a negative solver status; status 1 (fewer pairs
    than requested) warns only when verbose >= 1.
    """
    if device and not isinstance(A, (DeviceOperator, SparseDeviceOperator)):
        # scipy matrices (and a SuperLU or sparse matrix as preconditioner)
        # moved to the GPU as a general sparse operator
        if isinstance(A, np.ndarray):
            A = sp.csr_matrix(A)
        if not sp.issparse(A):
            raise ValueError("device=True needs A as a sparse matrix or ndarray")
        if B is not None and not (sp.issparse(B) or isinstance(B, np.ndarray)):
            raise ValueError("device=True needs B as a sparse matrix, ndarray or None")
        A = SparseDeviceOperator(A, B, preconditioner)
        B = None
        preconditioner = True if A.has_precond else None
    device = isinstance(A, (DeviceOperator, SparseDeviceOperator))
    if which not in _TARGET_MAP:
        raise ValueError(f"which must be one of {list(_TARGET_MAP.keys())}; got {which!r}")
    sigma = complex(sigma)
    target = _TARGET_MAP[which]

    if device:
        n = A.n
        if B is not None and B is not A:
            raise ValueError("with a device operator A, B comes from the same pencil (pass B=None)")
        if preconditioner is not None and preconditioner is not True and preconditioner is not A:
            raise ValueError("with a device operator A, preconditioner must be None, True or A")
        v0 = _start_vector(v0, n)
        out = _bbjd_c.jdqz_c(n, k, sigma, target, None, None,
                             None if preconditioner is None else True, tol=tol, maxiter=maxiter,
                             maxsub=maxsub, minsub=minsub, inner_maxiter=inner_maxiter,
                             inner_tol=inner_tol, verbose=verbose, v0=v0, device_op=A,
                             precision=precision, graphs=graphs, inner_form=inner_form)
        return _finish(out, k, verbose, return_info)

    # host operands: each becomes a function x -> y (complex128) by the first
    # matching kind of its table; dense arrays are taken as sparse matrices
    if isinstance(A, np.ndarray):
        A = sp.csr_matrix(A)
    n = A.shape[0]
    if tuple(A.shape) != (n, n):
        raise ValueError(f"A must be square; got shape {A.shape}")
    if isinstance(B, np.ndarray):
        B = sp.csr_matrix(B)
    matvec_A = _adapt(A, _OPERAND_KINDS, "A must be a sparse matrix, LinearOperator, or callable")
    matvec_B = (_identity if B is None else
                _adapt(B, _OPERAND_KINDS, "B must be a sparse matrix, LinearOperator, or callable"))
    precond_solve = (None if preconditioner is None else
                     _adapt(preconditioner, _PRECOND_KINDS,
                            "preconditioner must have a .solve(b) method or be callable"))
    v0 = _start_vector(v0, n)

    out = _bbjd_c.jdqz_c(n=n, nev=k, sigma=sigma, target=target, matvec_A=matvec_A,
                         matvec_B=matvec_B, precond_solve=precond_solve, tol=tol,
                         maxiter=maxiter, maxsub=maxsub, minsub=minsub,
                         inner_maxiter=inner_maxiter, inner_tol=inner_tol, verbose=verbose,
                         v0=v0, precision=precision, inner_form=inner_form)
    return _finish(out, k, verbose, return_info)


def _finish(out, k, verbose, return_info):
    evals, evecs, resid, nconv, iters, status = out
    if status < 0:
        msg = _STATUS_MSG.get(status, f"unknown error (status={status})")
        raise RuntimeError(f"JDQZ solver failed: {msg}")
    if status == 1 and verbose >= 1:
        import warnings
        warnings.warn(f"JDQZ: only {nconv}/{k} eigenvalues converged in {iters} iterations.",
                      RuntimeWarning, stacklevel=3)
    if return_info:
        return evals, evecs, {'nconv': nconv, 'iters': iters, 'status': status, 'residuals': resid}
    return evals, evecs
