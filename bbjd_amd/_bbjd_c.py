"""``jdqz_c``: the low-level entry of the reference's CPython extension
(reference src/_bbjd_c.c:128-259), over libbbjd.so.

Same positional/keyword arguments and return tuple
``(eigenvalues, eigenvectors, residuals, nconv, iters, status)`` with the
eigenvectors C-ordered n x nconv.  Python callables are called through ctypes
trampolines that pass a read-only view of the solver's vector, like the
reference's (src/_bbjd_c.c:43-123); an exception raised inside a callback
stops later callbacks and is re-raised as itself after the solve
(src/_bbjd_c.c:47, 212-216).  Deviation: a callback result of the wrong
length raises ValueError instead of being copied blindly (the reference
copies n elements unchecked, src/_bbjd_c.c:68).

Extension (keyword-only): ``device_op`` -- a ``DeviceOperator`` or a
``SparseDeviceOperator`` whose device-pointer callbacks replace the three
Python callables (no host round trip per apply); ``precision`` -- 'fp64' or 'mixed' (complex64 inner GMRES);
``qz_mode`` -- 'ritz' (ordered Schur form only where a restart may follow),
'schur' (every iteration, the reference's sequence) or 'device' (the
projected step on the GPU for k <= 64); ``graphs`` -- with
``device_op``, capture the inner GMRES steps into HIP graphs and replay
them (same kernels, same results); ``inner_form`` -- with precision
'mixed', 'reference' (the reference's GMRES, two-sided projections) or
'projected' (the projected preconditioner: one oblique projection per
step, bbjd_gpu_options.inner_form).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def _view(addr, n, writable):
    a = np.frombuffer((C.c_double * (2 * n)).from_address(addr), dtype=np.complex128)
    if not writable:
        a.flags.writeable = False
    return a


def _trampoline(fn, errbox):
    def cb(n, x, y, data):
        if errbox:
            return -1
        try:
            out = np.asarray(fn(_view(x, n, False)), dtype=np.complex128)
            if out.shape != (n,):
                out = out.reshape(-1)
                if out.shape != (n,):
                    raise ValueError(f"operator returned {out.size} entries, expected {n}")
            _view(y, n, True)[:] = out
            return 0
        except BaseException as e:          # noqa: BLE001 -- re-raised after the solve
            errbox.append(e)
            return -1
    return _lib.CALLBACK(cb)


def jdqz_c(n, nev, sigma, target, matvec_A, matvec_B, precond_solve, tol=1e-10, maxiter=200,
           maxsub=0, minsub=0, inner_maxiter=0, inner_tol=1e-2, verbose=0, v0=None, *,
           device_op=None, precision="fp64", return_stats=False, qz_mode="ritz", graphs=False,
           inner_form="reference"):
    lib = _lib.load()
    n, nev = int(n), int(nev)
    p = _lib.Params()
    lib.bbjd_params_init(C.byref(p))
    p.n, p.nev, p.sigma, p.target = n, nev, _lib.zc(sigma), int(target)
    p.tol, p.maxiter, p.maxsub, p.minsub = float(tol), int(maxiter), int(maxsub), int(minsub)
    p.inner_maxiter, p.inner_tol, p.verbose = int(inner_maxiter), float(inner_tol), int(verbose)
    v0_arr = None
    if v0 is not None:
        v0_arr = np.ascontiguousarray(v0, dtype=np.complex128)
        if v0_arr.ndim != 1 or v0_arr.shape[0] != n:
            raise ValueError(f"v0 must have shape ({n},)")
        p.v0 = v0_arr.ctypes.data
    opts = _lib.GpuOptions()
    lib.bbjd_gpu_options_init(C.byref(opts))
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    opts.precision = _lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64
    if qz_mode not in ("ritz", "schur", "device"):
        raise ValueError("qz_mode must be 'ritz', 'schur' or 'device'")
    opts.qz_mode = {"ritz": _lib.QZ_RITZ, "schur": _lib.QZ_SCHUR, "device": _lib.QZ_DEVICE}[qz_mode]
    opts.graphs = int(bool(graphs))
    if inner_form not in ("reference", "projected"):
        raise ValueError("inner_form must be 'reference' or 'projected'")
    opts.inner_form = _lib.INNER_PROJECTED if inner_form == "projected" else _lib.INNER_REFERENCE
    errbox: list = []
    keep = []
    if device_op is not None:
        # the operator names the C callbacks it provides; params takes the
        # three of the reference's contract, the options the optional ones
        opts.callback_space = _lib.CALLBACK_DEVICE
        p.precond = None
        for slot, name in device_op.device_callbacks(precision, precond_solve is not None).items():
            setattr(p if slot in ("matvec_A", "matvec_B", "precond") else opts, slot, _lib.fn_addr(name))
        p.op_data = device_op.handle
    else:
        cbA, cbB = _trampoline(matvec_A, errbox), _trampoline(matvec_B, errbox)
        keep += [cbA, cbB]
        p.matvec_A = C.cast(cbA, C.c_void_p).value
        p.matvec_B = C.cast(cbB, C.c_void_p).value
        if precond_solve is not None:
            cbP = _trampoline(precond_solve, errbox)
            keep.append(cbP)
            p.precond = C.cast(cbP, C.c_void_p).value
    ws = lib.bbjd_workspace_alloc_gpu(C.byref(p), C.byref(opts))
    if not ws:
        raise MemoryError("bbjd_workspace_alloc failed")
    res = _lib.Result()
    st = _lib.Stats()
    try:
        status = lib.bbjd_solve(C.byref(p), ws, C.byref(res))
        lib.bbjd_workspace_stats(ws, C.byref(st))
    finally:
        lib.bbjd_workspace_free(ws)
    if errbox:
        lib.bbjd_result_free(C.byref(res))
        raise errbox[0]
    nconv = int(res.nconv)
    evals = np.empty(nconv, dtype=np.complex128)
    evecs = np.empty((n, nconv), dtype=np.complex128)
    resid = np.empty(nconv, dtype=np.float64)
    if nconv > 0:
        evals[:] = np.frombuffer((C.c_double * (2 * nconv)).from_address(res.eigenvalues),
                                 dtype=np.complex128)
        ev = np.frombuffer((C.c_double * (2 * n * nconv)).from_address(res.eigenvectors),
                           dtype=np.complex128).reshape(nconv, n)
        evecs[:] = ev.T
        resid[:] = np.frombuffer((C.c_double * nconv).from_address(res.residuals), dtype=np.float64)
    iters = int(res.iters)
    lib.bbjd_result_free(C.byref(res))
    out = (evals, evecs, resid, nconv, iters, int(status))
    if return_stats:
        stats = {f: getattr(st, f) for f, _ in _lib.Stats._fields_}
        return out + (stats,)
    return out
