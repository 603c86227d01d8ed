"""ctypes binding of libbbjd.so -- the C ABI declared in include/bbjd.h,
include/bbjd_gpu.h and include/bbjd_ops.h.

The library is built in-tree (``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C bbjd_amd/csrc``).  Importing this module never
falls back to anything: when the shared object is missing, ``load()`` raises
ImportError.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbbjd.so")


class zcomplex(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


# int f(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *data)
CALLBACK = C.CFUNCTYPE(C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p)


class Params(C.Structure):
    """bbjd_params (field order of reference include/bbjd.h:114-150)."""
    _fields_ = [("n", C.c_int64), ("nev", C.c_int64), ("sigma", zcomplex), ("target", C.c_int),
                ("maxsub", C.c_int64), ("minsub", C.c_int64), ("tol", C.c_double),
                ("maxiter", C.c_int64), ("inner_maxiter", C.c_int64), ("inner_tol", C.c_double),
                ("v0", C.c_void_p), ("matvec_A", C.c_void_p), ("matvec_B", C.c_void_p),
                ("precond", C.c_void_p), ("op_data", C.c_void_p), ("verbose", C.c_int)]


class Result(C.Structure):
    """bbjd_result (reference include/bbjd.h:155-162)."""
    _fields_ = [("nconv", C.c_int64), ("eigenvalues", C.c_void_p), ("eigenvectors", C.c_void_p),
                ("residuals", C.c_void_p), ("iters", C.c_int64), ("status", C.c_int)]


QZ_RITZ, QZ_SCHUR, QZ_DEVICE = 0, 1, 2
INNER_REFERENCE, INNER_PROJECTED = 0, 1


class GpuOptions(C.Structure):
    _fields_ = [("callback_space", C.c_int), ("precision", C.c_int), ("stream", C.c_void_p),
                ("matvec_A_c64", C.c_void_p), ("matvec_B_c64", C.c_void_p),
                ("precond_c64", C.c_void_p), ("precond2", C.c_void_p), ("qz_mode", C.c_int),
                ("graphs", C.c_int), ("matvec_AB", C.c_void_p), ("matvec_AB_c64", C.c_void_p),
                ("matvec_ABr", C.c_void_p), ("inner_form", C.c_int), ("blocking_sync", C.c_int),
                ("inner_batcher", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("outer_iters", C.c_int64), ("lock_iters", C.c_int64), ("restarts", C.c_int64),
                ("inner_steps", C.c_int64), ("dgks_second_passes", C.c_int64),
                ("sum_k", C.c_int64), ("qz_seconds", C.c_double), ("solve_seconds", C.c_double),
                ("stream_bytes", C.c_double), ("restart_fallbacks", C.c_int64),
                ("graph_captures", C.c_int64), ("graph_launches", C.c_int64), ("qz_schur_ties", C.c_int64),
                ("qz_device_steps", C.c_int64), ("qz_device_fallbacks", C.c_int64),
                ("qz_device_accuracy", C.c_int64), ("inner_steps_discarded", C.c_int64),
                ("qz_harmonic_restarts", C.c_int64), ("inner_projected", C.c_int64),
                ("inner_projected_fallbacks", C.c_int64)]


class QzProblem(C.Structure):
    """bbjd_qz_problem: one projected step for bbjd_qz_device_batch (all
    pointers device addresses)"""
    _fields_ = [("k", C.c_int), ("ldm", C.c_int), ("MA", C.c_void_p), ("MB", C.c_void_p),
                ("nu0_re", C.c_double), ("nu0_im", C.c_double), ("mu0_re", C.c_double), ("mu0_im", C.c_double),
                ("sigma_re", C.c_double), ("sigma_im", C.c_double), ("target", C.c_int), ("nkeep", C.c_int),
                ("nconv", C.c_int), ("keep", C.c_int), ("conv", C.c_void_p), ("tol", C.c_double),
                ("alpha", C.c_void_p), ("beta", C.c_void_p), ("x", C.c_void_p), ("zr", C.c_void_p),
                ("zfull", C.c_void_p), ("tfull", C.c_void_p), ("info", C.c_void_p)]


class PencilDesc(C.Structure):
    _fields_ = [("n", C.c_int64), ("bs", C.c_int), ("hw", C.c_int), ("ntau", C.c_int),
                ("band", C.c_void_p), ("tau", C.c_void_p), ("bdiag", C.c_void_p)]


CALLBACK_HOST, CALLBACK_DEVICE = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
PC_NONE, PC_EXACT_LU, PC_BLOCK_JACOBI = 0, 1, 2

_lib = None


def _sig(lib, name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


def load():
    """Load libbbjd.so (once).  Raises ImportError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"bbjd_amd: {LIB_PATH} is not built; run __graft_entry__.build() "
                          "or make -C bbjd_amd/csrc")
    lib = C.CDLL(LIB_PATH)
    P, V, I64, D = C.POINTER(Params), C.c_void_p, C.c_int64, C.c_double
    _sig(lib, "bbjd_params_init", None, P)
    _sig(lib, "bbjd_workspace_alloc", V, P)
    _sig(lib, "bbjd_workspace_free", None, V)
    _sig(lib, "bbjd_solve", C.c_int, P, V, C.POINTER(Result))
    _sig(lib, "bbjd_result_free", None, C.POINTER(Result))
    _sig(lib, "bbjd_mgs_orthogonalize", D, I64, I64, V, V)
    _sig(lib, "bbjd_block_orthogonalize", D, I64, I64, V, V, V)
    _sig(lib, "bbjd_correction_gmres", C.c_int, I64, V, V, zcomplex, zcomplex, V, V, V, V, V, V,
         I64, D, V)
    _sig(lib, "bbjd_gpu_options_init", None, C.POINTER(GpuOptions))
    _sig(lib, "bbjd_workspace_bytes", C.c_size_t, P, C.POINTER(GpuOptions))
    _sig(lib, "bbjd_workspace_alloc_gpu", V, P, C.POINTER(GpuOptions))
    _sig(lib, "bbjd_callback_stream", V)
    _sig(lib, "bbjd_workspace_stats", C.c_int, V, C.POINTER(Stats))
    _sig(lib, "bbjd_dev_alloc", V, C.c_size_t)
    _sig(lib, "bbjd_dev_free", None, V)
    _sig(lib, "bbjd_dev_sync", C.c_int)
    _sig(lib, "bbjd_time_op", C.c_int, V, I64, V, V, V, C.c_int, C.POINTER(D))
    _sig(lib, "bbjd_kbench", C.c_int, C.c_int, I64, C.c_int, C.c_int, C.POINTER(D), C.POINTER(D))
    _sig(lib, "bbjd_set_tiles", C.c_int, C.c_int, C.c_int, C.c_int)
    _sig(lib, "bbjd_block_gram", C.c_int, I64, I64, C.c_void_p, I64, C.c_void_p, C.c_void_p)
    _sig(lib, "bbjd_set_fuse_blocks", None, C.c_int)
    _sig(lib, "bbjd_set_svdv_fused", None, C.c_int)
    _sig(lib, "bbjd_set_cgs_fused", None, C.c_int)
    _sig(lib, "bbjd_set_gram_blocks", None, C.c_int)
    _sig(lib, "bbjd_set_rotate_mfma", None, C.c_int)
    _sig(lib, "bbjd_set_bcr_group", None, C.c_int)
    _sig(lib, "bbjd_set_bcr_tail_rows", None, C.c_int)
    _sig(lib, "bbjd_set_bcr_border_fuse", None, C.c_int)
    _sig(lib, "bbjd_set_bcr_prefetch", None, C.c_int)
    _sig(lib, "bbjd_set_c64_direct_io", None, C.c_int)
    _sig(lib, "bbjd_set_cgs_mid_blocks", None, C.c_int)
    _sig(lib, "bbjd_set_tile_table", C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)
    _sig(lib, "bbjd_block_rotate", C.c_int, I64, I64, I64, C.c_void_p, C.c_void_p)
    _sig(lib, "bbjd_block_orthonormalize", D, I64, I64, C.c_void_p, C.c_void_p, C.c_void_p)
    _sig(lib, "bbjd_project_out_host", C.c_int, I64, I64, C.c_void_p, C.c_void_p)
    _sig(lib, "bbjd_expand_test_space", D, I64, I64, C.c_void_p, zcomplex, zcomplex, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_void_p, C.c_void_p)
    _sig(lib, "bbjd_memcpy_d2h", C.c_int, V, V, C.c_size_t)
    _sig(lib, "bbjd_memcpy_h2d", C.c_int, V, V, C.c_size_t)
    _sig(lib, "bbjd_pencil_generate", C.c_int, I64, C.c_int, C.c_int, C.c_int, C.c_uint64, D, V, V, V)
    _sig(lib, "bbjd_op_create", V, C.POINTER(PencilDesc), zcomplex, C.c_int, C.c_int, V)
    _sig(lib, "bbjd_op_free", None, V)
    _sig(lib, "bbjd_op_bytes", C.c_size_t, V)
    _sig(lib, "bbjd_op_precond2", C.c_int, I64, V, V, V, V, V)
    _sig(lib, "bbjd_op_matvec_AB", C.c_int, I64, V, V, V, V)
    _sig(lib, "bbjd_op_matvec_ABr", C.c_int, I64, V, V, V, zcomplex, zcomplex, V, V, V)
    _sig(lib, "bbjd_op_matvec_AB_c64", C.c_int, I64, V, V, V, V)
    for nm in ("bbjd_op_matvec_A", "bbjd_op_matvec_B", "bbjd_op_precond",
               "bbjd_op_matvec_A_c64", "bbjd_op_matvec_B_c64", "bbjd_op_precond_c64"):
        _sig(lib, nm, C.c_int, I64, V, V, V)
    _lib = lib
    return lib


class DeviceBuffer:
    """A device allocation made by the library (no other GPU runtime)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = load().bbjd_dev_alloc(self.nbytes)
        if not self.ptr:
            raise MemoryError(f"device allocation of {nbytes} bytes failed")

    @classmethod
    def from_array(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if load().bbjd_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes):
            raise RuntimeError("host-to-device copy failed")
        return b

    def to_array(self, dtype, count):
        import numpy as np
        a = np.empty(count, dtype=dtype)
        if load().bbjd_memcpy_d2h(a.ctypes.data, self.ptr, a.nbytes):
            raise RuntimeError("device-to-host copy failed")
        return a

    def free(self):
        if self.ptr:
            load().bbjd_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fn_addr(name: str) -> int:
    """Address of an exported C function (for params.matvec_A & co.)."""
    return C.cast(getattr(load(), name), C.c_void_p).value


def zc(z: complex) -> zcomplex:
    z = complex(z)
    return zcomplex(z.real, z.imag)
