"""Python access to the serving loop (include/bbjd_serve.h).

``serve(jobs, ...)`` runs a list of independent solves through the C
serving loop: W worker threads, each with its own stream and resident
workspace, admitting requests in arrival order (or largest first).
``Server`` keeps that loop, its threads and workspaces resident across
batches (``run(jobs)`` per batch, ``close()`` at the end).  A job
is a dict with ``op`` (a DeviceOperator already resident on the GPU) or
``pencil`` + ``pc`` (the server builds the operator at admission),
``sigma``, ``nev`` and optionally maxsub, minsub, maxiter, inner_maxiter,
tol, inner_tol, arrival (seconds), wave (arrival wave: its arrivals count
from the admission of the previous wave's last request), keep_vectors and v0
(a start vector of length n).
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
                ("stream_bytes", C.c_double), ("v0", C.c_void_p), ("wave", C.c_int), ("t_arrive", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("workers", C.c_int), ("hbm_budget", C.c_size_t), ("precision", C.c_int),
                ("realtime", C.c_int), ("sync_yield", C.c_int), ("graphs", C.c_int), ("order", C.c_int),
                ("pin", C.c_int), ("window", C.c_int), ("qz_mode", C.c_int), ("inner_form", C.c_int),
                ("progress", C.c_int), ("inner_batch", C.c_int), ("inner_batchers", C.c_int),
                ("inner_window_us", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("batches", C.c_long), ("jobs", C.c_long), ("steps", C.c_long), ("recorded", C.c_long),
                ("launches", C.c_long)]


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
    lib.bbjd_server_create.restype = C.c_void_p
    lib.bbjd_server_create.argtypes = [C.POINTER(Request), C.c_int, C.POINTER(ServeConfig)]
    lib.bbjd_server_run.restype = C.c_int
    lib.bbjd_server_run.argtypes = [C.c_void_p, C.POINTER(Request), C.c_int, C.POINTER(ServeConfig),
                                    C.POINTER(ServeReport)]
    lib.bbjd_server_destroy.argtypes = [C.c_void_p]
    lib.bbjd_server_workers.restype = C.c_int
    lib.bbjd_server_workers.argtypes = [C.c_void_p]
    lib.bbjd_server_workspace_bytes.restype = C.c_size_t
    lib.bbjd_server_workspace_bytes.argtypes = [C.c_void_p]
    lib.bbjd_server_batch_stats.argtypes = [C.c_void_p, C.POINTER(BatchStats)]
    lib.bbjd_device_free_bytes.restype = C.c_size_t
    lib.bbjd_device_free_bytes.argtypes = []
    lib._serve_ready = True


def device_init(sync_yield=True):
    lib = _lib.load()
    _setup(lib)
    return lib.bbjd_device_init(2 if sync_yield == "block" else int(bool(sync_yield)))


def _requests(jobs, keep_vectors):
    """bbjd_request array for the jobs, the host arrays it points into, and
    the output dicts"""
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
        q.wave = int(j.get("wave", 0))
        if j.get("v0") is not None:
            v0 = np.ascontiguousarray(j["v0"], dtype=np.complex128)
            if v0.shape != (n,):
                raise ValueError(f"v0 must have shape ({n},)")
            keep.append(v0)
            q.v0 = v0.ctypes.data
        o = {"eigenvalues": np.zeros(nev, np.complex128), "residuals": np.zeros(nev),
             "pair_residuals": np.zeros(nev)}
        if j.get("keep_vectors", keep_vectors):
            o["eigenvectors"] = np.zeros((nev, n), np.complex128)
            q.eigenvectors = o["eigenvectors"].ctypes.data
        q.eigenvalues = o["eigenvalues"].ctypes.data
        q.residuals = o["residuals"].ctypes.data
        q.pair_residuals = o["pair_residuals"].ctypes.data
        outs.append(o)
    return reqs, keep, outs


def _config(lib, workers=16, precision="fp64", realtime=True, hbm_budget=0, sync_yield=True, graphs=False,
            order="arrival", pin=False, window=8, qz_mode="ritz", inner_form="reference", progress=False,
            inner_batch=0, inner_batchers=2, inner_window_us=100.0):
    if order not in ("arrival", "largest", "window"):
        raise ValueError("order must be 'arrival', 'largest' or 'window'")
    if int(window) < 1:
        raise ValueError("window must be at least 1")
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
    modes = {"ritz": _lib.QZ_RITZ, "schur": _lib.QZ_SCHUR, "device": _lib.QZ_DEVICE}
    if qz_mode not in modes:
        raise ValueError("qz_mode must be 'ritz', 'schur' or 'device'")
    cfg.qz_mode = modes[qz_mode]
    forms = {"reference": _lib.INNER_REFERENCE, "projected": _lib.INNER_PROJECTED}
    if inner_form not in forms:
        raise ValueError("inner_form must be 'reference' or 'projected'")
    cfg.inner_form = forms[inner_form]
    cfg.progress = int(bool(progress))
    if not 0 <= int(inner_batch) <= 16 or not 1 <= int(inner_batchers) <= 8 or inner_window_us < 0:
        raise ValueError("inner_batch must be in [0, 16], inner_batchers in [1, 8], inner_window_us >= 0")
    cfg.inner_batch = int(inner_batch)
    cfg.inner_batchers = int(inner_batchers)
    cfg.inner_window_us = float(inner_window_us)
    return cfg


def _results(reqs, outs, keep_vectors):
    results = []
    for i in range(len(outs)):
        q = reqs[i]
        nc = int(q.nconv)
        o = outs[i]
        r = {"nconv": nc, "iters": int(q.iters), "status": int(q.status), "worker": int(q.worker),
             "arrival": q.arrival, "wave": q.wave, "t_arrive": q.t_arrive, "t_admit": q.t_admit, "t_start": q.t_start, "t_done": q.t_done,
             "qz_seconds": q.qz_seconds, "stream_bytes": q.stream_bytes,
             "eigenvalues": o["eigenvalues"][:nc].copy(),
             "residuals": o["residuals"][:nc].copy(),
             "pair_residuals": o["pair_residuals"][:nc].copy()}
        if "eigenvectors" in o:
            r["eigenvectors"] = o["eigenvectors"][:nc].T.copy()
        results.append(r)
    return results


def serve(jobs, workers=16, precision="fp64", realtime=True, hbm_budget=0, keep_vectors=False,
          sync_yield=True, graphs=False, order="arrival", pin=False, window=8, qz_mode="ritz",
          inner_form="reference", inner_batch=0, inner_batchers=2, inner_window_us=100.0):
    """Run jobs through the serving loop; returns (results, report).  graphs:
    capture each request's inner GMRES steps into HIP graphs and replay them.
    order: which arrived request a free worker takes, "arrival" (first come,
    first served), "largest" (largest estimated work: nev, then n) or
    "window" (the largest of the `window` longest-waiting ones).  qz_mode:
    where the projected step runs ("ritz": host, eigenvalues first;
    "schur": host, the reference's ordered Schur form every iteration;
    "device": on the GPU for subspaces up to 64).  inner_form: the mixed
    inner solve's form ("reference": the reference's GMRES with two-sided
    projections; "projected": the projected preconditioner, one oblique
    projection per step).  inner_batch > 0 (mixed precision): the workers'
    inner solves run in batches of up to inner_batch on inner_batchers
    threads, each Arnoldi step's kernels of a batch merged into shared
    launches; results are the same as without."""
    lib = _lib.load()
    _setup(lib)
    cfg = _config(lib, workers, precision, realtime, hbm_budget, sync_yield, graphs, order, pin, window, qz_mode,
                  inner_form, inner_batch=inner_batch, inner_batchers=inner_batchers,
                  inner_window_us=inner_window_us)
    reqs, keep, outs = _requests(jobs, keep_vectors)
    rep = ServeReport()
    rc = lib.bbjd_serve(reqs, len(jobs), C.byref(cfg), C.byref(rep))
    if rc:
        raise RuntimeError("bbjd_serve failed to set up (device memory?)")
    report = {f: getattr(rep, f) for f, _ in ServeReport._fields_}
    return _results(reqs, outs, keep_vectors), report


class Server:
    """A resident serving loop (bbjd_server_create): worker threads, streams
    and workspaces sized for the largest of ``capacity_jobs`` stay allocated
    until close(), and each run() serves a batch of jobs on them."""

    def __init__(self, capacity_jobs, workers=16, precision="fp64", hbm_budget=0, sync_yield=True,
                 graphs=False, pin=False, qz_mode="ritz", inner_form="reference", inner_batch=0,
                 inner_batchers=2, inner_window_us=100.0):
        lib = _lib.load()
        _setup(lib)
        self._lib = lib
        self.precision = precision
        self._cfg_kw = dict(workers=workers, precision=precision, hbm_budget=hbm_budget, sync_yield=sync_yield,
                            graphs=graphs, pin=pin, qz_mode=qz_mode, inner_form=inner_form,
                            inner_batch=inner_batch, inner_batchers=inner_batchers,
                            inner_window_us=inner_window_us)
        cfg = _config(lib, **self._cfg_kw)
        reqs, _keep, _outs = _requests(capacity_jobs, False)
        self._h = lib.bbjd_server_create(reqs, len(capacity_jobs), C.byref(cfg))
        if not self._h:
            raise RuntimeError("bbjd_server_create failed (device memory?)")

    @property
    def workers(self):
        return int(self._lib.bbjd_server_workers(self._h))

    @property
    def workspace_bytes(self):
        return int(self._lib.bbjd_server_workspace_bytes(self._h))

    def batch_stats(self):
        """counters of the grouped inner solves since creation: batches,
        jobs (inner solves), steps, recorded launches and launches made"""
        st = BatchStats()
        self._lib.bbjd_server_batch_stats(self._h, C.byref(st))
        return {f: int(getattr(st, f)) for f, _ in BatchStats._fields_}

    def run(self, jobs, realtime=True, order="arrival", window=8, keep_vectors=False, progress=False):
        """serve a batch; progress: a line on stderr as each arrival wave's
        last request completes"""
        if not self._h:
            raise RuntimeError("server is closed")
        cfg = _config(self._lib, realtime=realtime, order=order, window=window, progress=progress, **self._cfg_kw)
        reqs, keep, outs = _requests(jobs, keep_vectors)
        rep = ServeReport()
        if self._lib.bbjd_server_run(self._h, reqs, len(jobs), C.byref(cfg), C.byref(rep)):
            raise RuntimeError("bbjd_server_run failed")
        report = {f: getattr(rep, f) for f, _ in ServeReport._fields_}
        return _results(reqs, outs, keep_vectors), report

    def close(self):
        if getattr(self, "_h", None):
            self._lib.bbjd_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_free_bytes():
    lib = _lib.load()
    _setup(lib)
    return int(lib.bbjd_device_free_bytes())


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
