"""Pins of the oracle's building blocks by closed forms and invariants (for
the parts the reference's own tests leave open): Givens rotation, CGS2 with
the DGKS test, double projection, CPU pencil operators, the inner GMRES on
a tiny dense case, and the glibc rand() sequence the start vector uses."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from bbjd_amd.pencil import BandedTauPencil
from oracle import orc


def _zc(z):
    return orc.zcomplex(z.real, z.imag)


@pytest.mark.parametrize("f,g", [(1 + 2j, 3 - 1j), (0.0, 2 - 2j), (1e-3j, 0.0), (0, 0), (-5 + 0j, 1e-200 + 0j)])
def test_givens_closed_form(f, g):
    """reference src/jd_core.c:477-511: cs real >= 0, cs^2 + |sn|^2 = 1 and
    [cs sn; -conj(sn) cs] [f; g] = [r; 0]"""
    lib = orc.load_oracle()
    cs, sn, r = C.c_double(), orc.zcomplex(), orc.zcomplex()
    lib.orc_zlartg(_zc(complex(f)), _zc(complex(g)), C.byref(cs), C.byref(sn), C.byref(r))
    c, s, rr = cs.value, complex(sn.re, sn.im), complex(r.re, r.im)
    assert c >= 0
    assert abs(c * c + abs(s) ** 2 - 1) < 1e-14 or (f == 0 and g == 0)
    assert abs(c * f + s * g - rr) <= 1e-14 * max(1, abs(rr))
    assert abs(-np.conj(s) * f + c * g) <= 1e-14 * max(1, abs(f) + abs(g))
    if abs(g) < 1e-300:
        assert (c, s, rr) == (1.0, 0, complex(f))


def _orth(n, k, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))
    return np.asfortranarray(q)


def test_cgs2_invariants_and_dgks_branch():
    """reference src/jd_core.c:207-241: one pass leaves (I - QQ^H) v; the
    second pass runs iff the norm dropped below 0.7071 of the input's"""
    lib = orc.load_oracle()
    n, k = 500, 12
    Q = _orth(n, k, 1)
    rng = np.random.default_rng(2)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    # generic: ||(I-QQ^H)v|| ~ ||v|| -> one pass; equals the dense projection
    x = v.copy()
    h = np.zeros(k, np.complex128)
    nrm = lib.orc_block_orthogonalize(n, k, Q.ctypes.data, x.ctypes.data, h.ctypes.data)
    dense = v - Q @ (Q.conj().T @ v)
    assert np.linalg.norm(x - dense) <= 1e-13 * np.linalg.norm(v)
    assert abs(nrm - np.linalg.norm(x)) <= 1e-14 * nrm
    assert np.allclose(h, Q.conj().T @ v, atol=1e-13)       # single pass: h = Q^H v
    # nearly in span(Q): the second pass is taken -- h then holds the
    # second-pass coefficients Q^H (once-projected y), which are rounding-
    # sized, not the O(1) first-pass ones -- and orthogonality is restored
    c = rng.standard_normal(k)
    y = Q @ c + 1e-9 * v
    x = y.copy()
    nrm = lib.orc_block_orthogonalize(n, k, Q.ctypes.data, x.ctypes.data, h.ctypes.data)
    assert np.linalg.norm(h) <= 1e-12 * np.linalg.norm(y) < np.linalg.norm(c)
    assert np.linalg.norm(Q.conj().T @ x) <= 1e-14 * np.linalg.norm(y)
    assert abs(nrm - np.linalg.norm(x)) <= 1e-14 * np.linalg.norm(y)
    # k = 0: only the norm
    x = v.copy()
    assert abs(lib.orc_block_orthogonalize(n, 0, Q.ctypes.data, x.ctypes.data, h.ctypes.data)
               - np.linalg.norm(v)) < 1e-12


def test_project_out_two_passes():
    lib = orc.load_oracle()
    n, k = 300, 5
    Q = _orth(n, k, 3)
    v = np.random.default_rng(4).standard_normal(n) + 0j
    x = v.copy()
    lib.orc_project_out(n, k, Q.ctypes.data, x.ctypes.data)
    assert np.linalg.norm(Q.conj().T @ x) <= 1e-14 * np.linalg.norm(v)
    assert np.linalg.norm(x - (v - Q @ (Q.conj().T @ v))) <= 1e-13 * np.linalg.norm(v)


@pytest.mark.parametrize("n,bs,hw,ntau", [(400, 4, 1, 2), (512, 8, 2, 16), (96, 4, 2, 0), (808, 8, 1, 4)])
def test_cpu_pencil_operators(n, bs, hw, ntau):
    p = BandedTauPencil.generate(n, bs, hw, ntau, seed=7)
    A, B = p.to_scipy()
    sigma = 0.3 - 0.7j
    x = np.random.default_rng(0).standard_normal(n) + 1j * np.random.default_rng(1).standard_normal(n)
    lu = orc.CpuPencilOps(p, sigma, 1)
    assert np.allclose(lu.apply("orc_matvec_A", x), A @ x, rtol=0, atol=1e-12 * np.linalg.norm(A @ x))
    assert np.array_equal(lu.apply("orc_matvec_B", x), B @ x)
    M = (A - sigma * B).tocsc()
    y = lu.apply("orc_precond", x)
    assert np.linalg.norm(M @ y - x) <= 1e-11 * np.linalg.norm(x)
    assert np.linalg.norm(y - spla.spsolve(M, x)) <= 1e-9 * np.linalg.norm(y)
    bj = orc.CpuPencilOps(p, sigma, 2)
    z = bj.apply("orc_precond", x)
    Md = M.toarray()
    for b in range(0, n, bs):
        blk = Md[b:b + bs, b:b + bs]
        assert np.allclose(blk @ z[b:b + bs], x[b:b + bs], atol=1e-11 * np.linalg.norm(x[b:b + bs]))


def test_gmres_solves_projected_equation_when_complete():
    """reference src/jd_core.c:568-727 with as many steps as the projected
    space has dimensions and a tiny tolerance: the result solves
    M^-1 P_L (tb A - ta B) P_R t = -M^-1 P_L r and is orthogonal to Q_R"""
    n = 40
    p = BandedTauPencil.generate(n, 4, 1, 2, seed=3)
    A, B = p.to_scipy()
    A, B = A.toarray(), B.toarray()
    ops = orc.CpuPencilOps(p, 0.2 + 0.1j, 1)
    rng = np.random.default_rng(5)
    qr_ = np.linalg.qr(rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2)))[0]
    ql_ = np.linalg.qr(rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2)))[0]
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ta, tb = 0.4 + 0.3j, 1.1 - 0.2j
    z = orc.gmres(n, qr_, ql_, ta, tb, r, ops, maxiter=n, tol=1e-15)
    PL = np.eye(n) - ql_ @ ql_.conj().T
    PR = np.eye(n) - qr_ @ qr_.conj().T
    Minv = np.linalg.inv(A - (0.2 + 0.1j) * B)
    lhs = Minv @ PL @ (tb * A - ta * B) @ PR @ z
    rhs = -Minv @ PL @ r
    assert np.linalg.norm(lhs - rhs) <= 1e-8 * np.linalg.norm(rhs)
    assert np.linalg.norm(qr_.conj().T @ z) <= 1e-12 * np.linalg.norm(z)


def test_glibc_rand_sequence_reproduced_by_random_r():
    """the engine reproduces srand(42); rand() (reference src/jd_core.c:
    793-799) with random_r on a private 128-byte state: same sequence"""
    libc = C.CDLL("libc.so.6")
    libc.srand(42)
    seq = [libc.rand() for _ in range(1000)]

    class RandomData(C.Structure):
        _fields_ = [("fptr", C.c_void_p), ("rptr", C.c_void_p), ("state", C.c_void_p),
                    ("rand_type", C.c_int), ("rand_deg", C.c_int), ("rand_sep", C.c_int),
                    ("end_ptr", C.c_void_p)]
    rd = RandomData()
    buf = (C.c_int32 * 32)()
    assert libc.initstate_r(42, buf, 128, C.byref(rd)) == 0
    out = C.c_int32()
    mine = []
    for _ in range(1000):
        libc.random_r(C.byref(rd), C.byref(out))
        mine.append(out.value)
    assert mine == seq
