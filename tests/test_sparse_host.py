"""Host logic of the general sparse operators (include/bbjd_sparse.h), no
GPU: the factor convention the device solve implements, the level sets of
its triangular solves, and the validation that refuses bad input before any
device call."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve_triangular, splu

from bbjd_amd import _lib
from bbjd_amd import sparse as bsp


def _random(n, density, seed, shift=3.0):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=seed, format="csr").astype(np.complex128)
    M.data += 0.3j * rng.standard_normal(M.nnz)
    return (M + shift * sp.eye(n)).tocsr()


def test_superlu_factors_give_the_solve_the_device_applies():
    """M^-1 b = Pc U^-1 L^-1 Pr b with (Pr b)[perm_r[i]] = b[i] and
    (Pc y)[i] = y[perm_c[i]] -- the sequence bbjd_sparse_precond runs (gather,
    two triangular solves, gather) -- equals SuperLU's own solve"""
    M = _random(400, 0.02, 1)
    lu = splu(M.tocsc())
    rng = np.random.default_rng(2)
    b = rng.standard_normal(400) + 1j * rng.standard_normal(400)
    t = np.empty_like(b)
    t[lu.perm_r] = b
    z = spsolve_triangular(sp.csr_matrix(lu.L), t, lower=True)
    w = spsolve_triangular(sp.csr_matrix(lu.U), z, lower=False)
    assert np.linalg.norm(w[lu.perm_c] - lu.solve(b)) <= 1e-12 * np.linalg.norm(lu.solve(b))
    assert np.allclose(lu.L.diagonal(), 1.0)          # unit lower factor, diagonal stored


def _levels_brute(F, lower):
    F = sp.csr_matrix(F)
    n = F.shape[0]
    lev = np.zeros(n, int)
    rows = range(n) if lower else range(n - 1, -1, -1)
    for i in rows:
        deps = [j for j in F.indices[F.indptr[i]:F.indptr[i + 1]] if j != i]
        lev[i] = 1 + max(lev[j] for j in deps) if deps else 0
    return lev


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_levels_match_the_recurrence(seed):
    """bbjd_sparse_levels on the factors of random sparse LUs against the
    defining recurrence level(i) = 1 + max level of the rows it needs; rows
    of one level never depend on each other"""
    M = _random(300, 0.015, seed)
    lu = splu(M.tocsc())
    for F, lower in ((lu.L, 1), (lu.U, 0)):
        nlev, lev = bsp.levels(F, lower)
        ref = _levels_brute(F, lower)
        assert nlev == ref.max() + 1
        assert np.array_equal(lev, ref)
        Fr = sp.csr_matrix(F)
        for i in range(300):
            deps = [j for j in Fr.indices[Fr.indptr[i]:Fr.indptr[i + 1]] if j != i]
            assert all(lev[j] < lev[i] for j in deps)


def test_levels_special_shapes():
    n = 50
    assert bsp.levels(sp.eye(n), 1)[0] == 1                                    # diagonal: one level
    T = sp.diags([np.ones(n - 1), 2 * np.ones(n)], [-1, 0])
    nlev, lev = bsp.levels(T, 1)                                                # bidiagonal: a chain
    assert nlev == n and np.array_equal(lev, np.arange(n))
    assert bsp.levels(T.T, 1)[0] == -1                                          # an entry above the diagonal
    assert bsp.levels(T, 0)[0] == -1


def test_create_refuses_bad_factors_before_any_device_call():
    """shapes, a factor on the wrong side of the diagonal or without its
    diagonal entry, a permutation out of range: NULL, decided on the host"""
    lib = _lib.load()
    bsp._setup(lib)
    n = 20
    keep = []
    A = bsp._csr(_random(n, 0.2, 4), n, keep)
    lu = splu(_random(n, 0.2, 5).tocsc())
    pr = np.ascontiguousarray(lu.perm_r, np.int32)
    pc = np.ascontiguousarray(lu.perm_c, np.int32)
    bad_pc = pc.copy()
    bad_pc[0] = n                                                               # out of range
    dup_pr = pr.copy()
    dup_pr[1] = dup_pr[0]                                                       # not a permutation
    Lc, Uc = bsp._csr(lu.L, n, keep), bsp._csr(lu.U, n, keep)
    no_diag = sp.csr_matrix(lu.U)
    no_diag.setdiag(0)
    no_diag.eliminate_zeros()
    for L, U, p_r, p_c in ((Uc, Uc, pr, pc), (Lc, Lc, pr, pc), (Lc, bsp._csr(no_diag, n, keep), pr, pc),
                           (Lc, Uc, pr, bad_pc), (Lc, Uc, dup_pr, pc)):
        f = bsp.SparseLU(L, U, p_r.ctypes.data, p_c.ctypes.data)
        assert not lib.bbjd_sparse_op_create(C.byref(A), None, C.byref(f), None)
    wrong = bsp._csr(_random(n + 1, 0.2, 6), n + 1, keep)
    assert not lib.bbjd_sparse_op_create(C.byref(A), C.byref(wrong), None, None)


def test_dense_rows_pinned_out_of_the_sparse_lu():
    """the bordered factorization of a banded pencil with 4 dense rows
    (the benchmark family as a scipy matrix): the detection finds exactly
    the dense rows, M11's LU stays sparse (the whole matrix's SuperLU fills
    U with O(n^2) entries), and the Schur-complement solve reproduces a
    dense solve to 1e-12; a matrix without dense rows is left alone"""
    import scipy.sparse as sp
    from bbjd_amd import BandedTauPencil
    from bbjd_amd.sparse import bordered_solve, dense_rows, pinned_factors
    p = BandedTauPencil.generate(1200, 8, 1, 4, seed=3)
    A, B = p.to_scipy()
    M = (A - (0.5 + 0.5j) * B).tocsr()
    rows = dense_rows(M)
    assert rows.tolist() == list(range(p.n1, p.n))
    f = pinned_factors(M, rows)
    assert f["lu11"].U.nnz < 20 * M.nnz
    rng = np.random.default_rng(0)
    b = rng.standard_normal(p.n) + 1j * rng.standard_normal(p.n)
    ref = np.linalg.solve(M.toarray(), b)
    assert np.linalg.norm(bordered_solve(f, b) - ref) <= 1e-12 * np.linalg.norm(ref)
    assert dense_rows(sp.random(500, 500, density=0.02, random_state=1, format="csr") + sp.eye(500)).size == 0
