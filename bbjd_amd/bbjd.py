"""``jdqz``: scipy-eigs-style Jacobi-Davidson QZ on one MI355X.

Mirrors the reference wrapper (kyle-augustson/BBJD bbjd.py:59-277): same
signature, defaults, input handling, error mapping and return values, over
the GPU engine (libbbjd.so).

    evals, evecs = jdqz(A, B, k=10, sigma=0.5, preconditioner=lu)

A and B may be dense arrays (-> CSR), scipy sparse matrices (CSC matvec),
objects with .matvec, or callables (reference bbjd.py:161-205); the
preconditioner an object with .solve or a callable (bbjd.py:208-219).  These
run on the host and the engine stages each vector to/from HBM.

Device fast path (extension): pass a ``DeviceOperator`` as A (B None) and
the same or another DeviceOperator, or True, as ``preconditioner``; then the
matvecs and the preconditioner run on the GPU as well and no n-vector
crosses PCIe during the iteration.  ``precision='mixed'`` runs the inner
GMRES (inner_maxiter > 0) in complex64 with FP32 accumulation, and
``graphs=True`` replays its steps from captured HIP graphs.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _bbjd_c
from .pencil import DeviceOperator

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


def jdqz(A, B=None, k: int = 6, sigma: complex = 0.0, which: str = 'nearest', tol: float = 1e-10,
         maxiter: int = 200, maxsub: int = 0, minsub: int = 0, inner_maxiter: int = 0,
         inner_tol: float = 1e-2, preconditioner=None, v0=None, verbose: int = 0,
         return_info: bool = False, precision: str = 'fp64', graphs: bool = False):
    """Jacobi-Davidson QZ for A x = lambda B x, eigenvalues nearest sigma
    (or by `which`).

    maxsub = 0 picks min(max(3k + 10, 40), n) and minsub = 0 picks
    max(2k + 5, 20) (the C defaults, reference src/jd_core.c:95-98, 750-752).
    Returns (eigenvalues, eigenvectors) or, with return_info, also
    {'nconv', 'iters', 'status', 'residuals'}.  Raises ValueError on invalid
    input and RuntimeError on a negative solver status; status 1 (fewer pairs
    than requested) warns only when verbose >= 1.
    """
    device = isinstance(A, DeviceOperator)
    if which not in _TARGET_MAP:
        raise ValueError(f"which must be one of {list(_TARGET_MAP.keys())}; got {which!r}")
    sigma = complex(sigma)
    target = _TARGET_MAP[which]

    if device:
        n = A.n
        if B is not None and B is not A:
            raise ValueError("with a DeviceOperator A, B comes from the same pencil (pass B=None)")
        if preconditioner is not None and preconditioner is not True and preconditioner is not A:
            raise ValueError("with a DeviceOperator A, preconditioner must be None, True or A")
        if v0 is not None:
            v0 = np.ascontiguousarray(v0, dtype=np.complex128)
            if v0.shape != (n,):
                raise ValueError(f"v0 must have shape ({n},); got {v0.shape}")
        out = _bbjd_c.jdqz_c(n, k, sigma, target, None, None,
                             None if preconditioner is None else True, tol=tol, maxiter=maxiter,
                             maxsub=maxsub, minsub=minsub, inner_maxiter=inner_maxiter,
                             inner_tol=inner_tol, verbose=verbose, v0=v0, device_op=A,
                             precision=precision, graphs=graphs)
        return _finish(out, k, verbose, return_info)

    if isinstance(A, np.ndarray):
        A = sp.csr_matrix(A)
    n = A.shape[0]
    if A.shape != (n, n):
        raise ValueError(f"A must be square; got shape {A.shape}")
    if B is None:
        B_op = sp.eye(n, dtype=np.complex128, format='csr')
    elif isinstance(B, np.ndarray):
        B_op = sp.csr_matrix(B)
    else:
        B_op = B

    if sp.issparse(A):
        A_csc = A.astype(np.complex128).tocsc()

        def matvec_A(x):
            return A_csc @ x
    elif hasattr(A, 'matvec'):
        def matvec_A(x):
            return np.asarray(A.matvec(x), dtype=np.complex128)
    elif callable(A):
        matvec_A = A
    else:
        raise ValueError("A must be a sparse matrix, LinearOperator, or callable")

    if sp.issparse(B_op):
        B_csc = B_op.astype(np.complex128).tocsc()

        def matvec_B(x):
            return B_csc @ x
    elif hasattr(B_op, 'matvec'):
        def matvec_B(x):
            return np.asarray(B_op.matvec(x), dtype=np.complex128)
    elif callable(B_op):
        matvec_B = B_op
    else:
        raise ValueError("B must be a sparse matrix, LinearOperator, or callable")

    if preconditioner is not None:
        if hasattr(preconditioner, 'solve'):
            pc = preconditioner

            def precond_solve(x):
                return np.asarray(pc.solve(x), dtype=np.complex128)
        elif callable(preconditioner):
            precond_solve = preconditioner
        else:
            raise ValueError("preconditioner must have a .solve(b) method or be callable")
    else:
        precond_solve = None

    if v0 is not None:
        v0 = np.ascontiguousarray(v0, dtype=np.complex128)
        if v0.shape != (n,):
            raise ValueError(f"v0 must have shape ({n},); got {v0.shape}")

    out = _bbjd_c.jdqz_c(n=n, nev=k, sigma=sigma, target=target, matvec_A=matvec_A,
                         matvec_B=matvec_B, precond_solve=precond_solve, tol=tol,
                         maxiter=maxiter, maxsub=maxsub, minsub=minsub,
                         inner_maxiter=inner_maxiter, inner_tol=inner_tol, verbose=verbose,
                         v0=v0, precision=precision)
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
