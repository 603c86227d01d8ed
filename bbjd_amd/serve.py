"""Python access to the serving loop (include/bbjd_serve.h).

``serve(jobs, ...)`` runs a list of independent solves through the C
serving loop: W worker threads, each with its own stream and resident
workspace, admitting requests in arrival order (or largest first).  A job
is a dict with ``op`` (a DeviceOperator already resident on the GPU) or
``pencil`` + ``pc`` (the server builds the operator at admission),
``sigma``, ``nev`` and optionally maxsub, minsub, maxiter, inner_maxiter,
tol, inner_tol, arrival (seconds) and v0 (a start vector of length n).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .pencil import _PC


class Request(C.Structure):
    _fields_ = [("op", C.c_void_p), ("pencil", C.c_void_p), ("pc", C.c_int), ("sigma", _lib.zcomplex),
                ("nev", C.c_int64), ("maxsub", C.c_int64), ("minsub", C.c_int64),
                ("maxiter", C.c_int64), ("inner_maxiter", C.c_int64), ("tol", C.c_double),
                ("inner_tol", C.c_double), ("arrival", C.c_double), ("eigenvalues", C.c_void_p),
                ("residuals", C.c_void_p), ("pair_residuals", C.c_void_p),
                ("eigenvectors", C.c_void_p), ("nconv", C.c_int64), ("iters", C.c_int64),
                ("status", C.c_int), ("worker", C.c_int), ("t_admit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double), ("qz_seconds", C.c_double),
                ("stream_bytes", C.c_double), ("v0", C.c_void_p)]


class ServeConfig(C.Structure):
    _fields_ = [("workers", C.c_int), ("hbm_budget", C.c_size_t), ("precision", C.c_int),
                ("realtime", C.c_int), ("sync_yield", C.c_int), ("graphs", C.c_int), ("order", C.c_int),
                ("pin", C.c_int), ("window", C.c_int)]


class ServeReport(C.Structure):
    _fields_ = [("wall_seconds", C.c_double), ("pairs", C.c_int64), ("failed", C.c_int64),
                ("workers", C.c_int), ("workspace_bytes", C.c_size_t), ("peak_bytes", C.c_size_t),
                ("peak_concurrency", C.c_int)]


def _setup(lib):
    if getattr(lib, "_serve_ready", False):
        return
    lib.bbjd_serve.restype = C.c_int
    lib.bbjd_serve.argtypes = [C.POINTER(Request), C.c_int, C.POINTER(ServeConfig),
                               C.POINTER(ServeReport)]
    lib.bbjd_serve_config_init.argtypes = [C.POINTER(ServeConfig)]
    lib.bbjd_device_init.argtypes = [C.c_int]
    lib.bbjd_device_init.restype = C.c_int
    lib.bbjd_ktime_enable.argtypes = [C.c_int]
    lib.bbjd_ktime_harvest.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.bbjd_ktime_harvest.restype = C.c_int
    lib.bbjd_op_bytes_for.argtypes = [C.POINTER(_lib.PencilDesc), C.c_int, C.c_int]
    lib.bbjd_op_bytes_for.restype = C.c_size_t
    lib._serve_ready = True


def device_init(sync_yield=True):
    lib = _lib.load()
    _setup(lib)
    return lib.bbjd_device_init(2 if sync_yield == "block" else int(bool(sync_yield)))


def serve(jobs, workers=16, precision="fp64", realtime=True, hbm_budget=0, keep_vectors=False,
          sync_yield=True, graphs=False, order="arrival", pin=False, window=8):
    """Run jobs through the serving loop; returns (results, report).  graphs:
    capture each request's inner GMRES steps into HIP graphs and replay them.
    order: which arrived request a free worker takes, "arrival" (first come,
    first served), "largest" (largest estimated work: nev, then n) or
    "window" (the largest of the `window` longest-waiting ones)."""
    if order not in ("arrival", "largest", "window"):
        raise ValueError("order must be 'arrival', 'largest' or 'window'")
    if int(window) < 1:
        raise ValueError("window must be at least 1")
    lib = _lib.load()
    _setup(lib)
    nreq = len(jobs)
    reqs = (Request * nreq)()
    keep = []
    outs = []
    for i, j in enumerate(jobs):
        q = reqs[i]
        nev = int(j["nev"])
        if j.get("op") is not None:
            q.op = j["op"].handle
            n = j["op"].n
        else:
            d = j["pencil"].desc()
            keep.append(d)
            q.pencil = C.addressof(d)
            q.pc = _PC[j.get("pc", "exact")]
            n = j["pencil"].n
        q.sigma = _lib.zc(j["sigma"])
        q.nev = nev
        q.maxsub = int(j.get("maxsub", 0))
        q.minsub = int(j.get("minsub", 0))
        q.maxiter = int(j.get("maxiter", 200))
        q.inner_maxiter = int(j.get("inner_maxiter", 0))
        q.tol = float(j.get("tol", 1e-10))
        q.inner_tol = float(j.get("inner_tol", 1e-2))
        q.arrival = float(j.get("arrival", 0.0))
        if j.get("v0") is not None:
            v0 = np.ascontiguousarray(j["v0"], dtype=np.complex128)
            if v0.shape != (n,):
                raise ValueError(f"v0 must have shape ({n},)")
            keep.append(v0)
            q.v0 = v0.ctypes.data
        o = {"eigenvalues": np.zeros(nev, np.complex128), "residuals": np.zeros(nev),
             "pair_residuals": np.zeros(nev)}
        if keep_vectors:
            o["eigenvectors"] = np.zeros((nev, n), np.complex128)
            q.eigenvectors = o["eigenvectors"].ctypes.data
        q.eigenvalues = o["eigenvalues"].ctypes.data
        q.residuals = o["residuals"].ctypes.data
        q.pair_residuals = o["pair_residuals"].ctypes.data
        outs.append(o)
    cfg = ServeConfig()
    lib.bbjd_serve_config_init(C.byref(cfg))
    cfg.workers = int(workers)
    cfg.hbm_budget = int(hbm_budget)
    cfg.precision = _lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64
    cfg.realtime = int(bool(realtime))
    cfg.sync_yield = 2 if sync_yield == "block" else int(bool(sync_yield))
    cfg.graphs = int(bool(graphs))
    cfg.order = {"arrival": 0, "largest": 1, "window": 2}[order]
    cfg.window = int(window)
    cfg.pin = int(bool(pin))
    rep = ServeReport()
    rc = lib.bbjd_serve(reqs, nreq, C.byref(cfg), C.byref(rep))
    if rc:
        raise RuntimeError("bbjd_serve failed to set up (device memory?)")
    results = []
    for i in range(nreq):
        q = reqs[i]
        nc = int(q.nconv)
        o = outs[i]
        r = {"nconv": nc, "iters": int(q.iters), "status": int(q.status), "worker": int(q.worker),
             "arrival": q.arrival, "t_admit": q.t_admit, "t_start": q.t_start, "t_done": q.t_done,
             "qz_seconds": q.qz_seconds, "stream_bytes": q.stream_bytes,
             "eigenvalues": o["eigenvalues"][:nc].copy(),
             "residuals": o["residuals"][:nc].copy(),
             "pair_residuals": o["pair_residuals"][:nc].copy()}
        if keep_vectors:
            r["eigenvectors"] = o["eigenvectors"][:nc].T.copy()
        results.append(r)
    report = {f: getattr(rep, f) for f, _ in ServeReport._fields_}
    return results, report


def ktime_enable(on=True):
    lib = _lib.load()
    _setup(lib)
    lib.bbjd_ktime_enable(int(on))


def ktime_harvest():
    lib = _lib.load()
    _setup(lib)
    ms, by = C.c_double(0), C.c_double(0)
    cnt = lib.bbjd_ktime_harvest(C.byref(ms), C.byref(by))
    return cnt, ms.value, by.value
