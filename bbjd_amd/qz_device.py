"""The projected step on the device, many problems per launch
(bbjd_qz_device_batch, include/bbjd_gpu.h).

``step_batch(problems)`` lays every problem's inputs and outputs out in one
device buffer, launches one wavefront per problem, and returns per problem
the values (alpha, beta), the target position, the number selected, the
Ritz coefficients and, on request, the Schur form and the restart rotation.
The engine calls the same entry point for its own step (qz_mode
``"device"``); this module serves tests and timing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

STATUS = ("ok", "ambiguous", "not_triangular", "accuracy", "no_convergence", "unsupported")


def _setup(lib):
    if getattr(lib, "_qzd_ready", False):
        return
    lib.bbjd_qz_device_batch.restype = C.c_int
    lib.bbjd_qz_device_batch.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    lib._qzd_ready = True


class Batch:
    """Problems laid out on the device once; launch() runs them all in one
    bbjd_qz_device_batch call, results() reads them back."""

    def __init__(self, problems, full=False):
        lib = _lib.load()
        _setup(lib)
        self.lib = lib
        self.items = []
        off = 0

        def take(nbytes):
            nonlocal off
            o = off
            off += (int(nbytes) + 255) // 256 * 256
            return o

        for q in problems:
            MA = np.asfortranarray(q["MA"], np.complex128)
            MB = np.asfortranarray(q["MB"], np.complex128)
            k = MA.shape[0]
            if MA.shape != (k, k) or MB.shape != (k, k) or not 1 <= k <= 64:
                raise ValueError("each problem needs square MA, MB of the same size, 1 <= k <= 64")
            conv = np.ascontiguousarray(np.asarray(q.get("conv", ()), np.complex128))
            keep = int(q.get("keep", 0))
            it = dict(k=k, MA=MA, MB=MB, conv=conv, keep=keep, q=q,
                      o_ma=take(MA.nbytes), o_mb=take(MB.nbytes), o_cv=take(max(conv.nbytes, 16)),
                      o_al=take(16 * k), o_be=take(16 * k), o_x=take(16 * k), o_info=take(16),
                      o_zr=take(16 * k * keep) if keep else None,
                      o_z=take(16 * k * k) if full else None, o_t=take(16 * k * k) if full else None)
            self.items.append(it)
        self.nbytes = max(off, 256)
        self.buf = _lib.DeviceBuffer(self.nbytes)
        host = np.zeros(self.nbytes, np.uint8)
        base = self.buf.ptr
        probs = (_lib.QzProblem * len(self.items))()
        for i, it in enumerate(self.items):
            q = it["q"]
            for key, arr in (("o_ma", it["MA"]), ("o_mb", it["MB"]), ("o_cv", it["conv"])):
                raw = np.frombuffer(np.asfortranarray(arr).tobytes(order="F"), np.uint8)
                host[it[key]: it[key] + raw.size] = raw
            sigma = complex(q.get("sigma", 0.0))
            nu0 = complex(q.get("nu0", 1 / np.sqrt(1 + abs(sigma) ** 2)))
            mu0 = complex(q.get("mu0", -sigma * nu0))
            p = probs[i]
            p.k, p.ldm = it["k"], it["k"]
            p.MA, p.MB = base + it["o_ma"], base + it["o_mb"]
            p.nu0_re, p.nu0_im, p.mu0_re, p.mu0_im = nu0.real, nu0.imag, mu0.real, mu0.imag
            p.sigma_re, p.sigma_im = sigma.real, sigma.imag
            p.target, p.nkeep = int(q.get("target", 0)), int(q.get("nkeep", 1))
            p.nconv, p.keep = it["conv"].size, it["keep"]
            p.conv = base + it["o_cv"]
            p.tol = float(q.get("tol", 1e-10))
            p.alpha, p.beta, p.x = base + it["o_al"], base + it["o_be"], base + it["o_x"]
            p.zr = base + it["o_zr"] if it["keep"] else None
            p.zfull = base + it["o_z"] if full else None
            p.tfull = base + it["o_t"] if full else None
            p.info = base + it["o_info"]
        if lib.bbjd_memcpy_h2d(base, host.ctypes.data, self.nbytes):
            raise RuntimeError("host-to-device copy failed")
        self.probs = _lib.DeviceBuffer.from_array(np.frombuffer(bytes(probs), np.uint8))
        self.full = full

    def launch(self, stream=None):
        if self.lib.bbjd_qz_device_batch(self.probs.ptr, len(self.items), stream):
            raise RuntimeError("bbjd_qz_device_batch failed")

    def results(self):
        if self.lib.bbjd_dev_sync():
            raise RuntimeError("device error after bbjd_qz_device_batch")
        raw = self.buf.to_array(np.uint8, self.nbytes)
        out = []

        def arr(o, count, dtype=np.complex128):
            return raw[o: o + count * np.dtype(dtype).itemsize].view(dtype).copy()

        for it in self.items:
            k = it["k"]
            info = arr(it["o_info"], 4, np.int32)
            r = dict(alpha=arr(it["o_al"], k), beta=arr(it["o_be"], k), x=arr(it["o_x"], k),
                     status=int(info[0]), p=int(info[1]), nsel=int(info[2]), its=int(info[3]))
            if it["keep"]:
                r["zr"] = arr(it["o_zr"], k * it["keep"]).reshape(k, it["keep"])
            if self.full:
                r["Z"] = arr(it["o_z"], k * k).reshape(k, k, order="F")
                r["T"] = arr(it["o_t"], k * k).reshape(k, k, order="F")
            out.append(r)
        return out

    def free(self):
        self.buf.free()
        self.probs.free()


def step_batch(problems, full=False):
    """every problem's projected step in one launch (see the module
    docstring); each problem: MA, MB (k x k, k <= 64), sigma, and optionally
    nu0, mu0 (default the harmonic test space's 1/sqrt(1+|sigma|^2) and
    -sigma nu0), target (bbjd_target), nkeep, conv (converged values), tol,
    keep (columns of the restart rotation)"""
    b = Batch(problems, full)
    try:
        b.launch()
        return b.results()
    finally:
        b.free()
