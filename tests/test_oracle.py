"""The oracle (CPU restatement of the reference solver, oracle/orc_jdqz.c)
pinned against the reference itself: the golden fixtures written by the
reference solver compiled from its own sources (tests/golden/
reference_runs.json, see make_golden.py) and, where the build container has
the reference, a live run of it.  The restatement uses the same LAPACK/BLAS
calls in the same order, so it must reproduce the reference's iteration
count and eigenvalues to the last bits."""
import json
import os

import numpy as np
import pytest

from oracle import orc
from tests.golden.make_golden import cases

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "reference_runs.json")))["cases"]
CASES = dict(cases())


def _evals(rec):
    return np.array([complex(a, b) for a, b in rec["evals"]])


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_oracle_reproduces_reference_fixture(name):
    orc.load_oracle().orc_set_threads(1)
    kw = CASES[name]()
    kw.pop("_keep", None)
    kw.pop("_pencil", None)
    kw.pop("_mats", None)
    r = orc.solve("oracle", **kw)
    g = GOLDEN[name]
    assert r["status"] == g["status"]
    assert r["nconv"] == g["nconv"]
    assert r["iters"] == g["iters"]
    ge = _evals(g)
    assert np.max(np.abs(r["evals"] - ge) / np.maximum(np.abs(ge), 1e-300), initial=0) <= 1e-13
    assert np.all(np.asarray(r["resid"]) < kw.get("tol", 1e-10))


@pytest.mark.skipif(not orc.have_ref(), reason="reference sources/build not present")
@pytest.mark.parametrize("name", ["diagonal_5_nearest_sigma0", "random_gmres_inner20", "config1_n2000"])
def test_fixture_is_current_with_the_reference_build(name):
    orc.load_oracle().orc_set_threads(1)
    kw = CASES[name]()
    kw.pop("_keep", None)
    kw.pop("_pencil", None)
    kw.pop("_mats", None)
    r = orc.solve("ref", **kw)
    g = GOLDEN[name]
    assert (r["nconv"], r["iters"], r["status"]) == (g["nconv"], g["iters"], g["status"])
    assert np.allclose(r["evals"], _evals(g), rtol=1e-13, atol=0)


# ---- the fixtures against what the mathematics fixes ---------------------

def test_fixture_known_answers():
    """known spectra of the reference's own tests (tests/test_bbjd.py:162-174,
    335-347, 509-525, 527-543, 444-457)"""
    e = np.sort(_evals(GOLDEN["diagonal_5_nearest_sigma0"]).real)
    assert np.allclose(e, [0.1, 0.2, 0.3, 0.4, 0.5], atol=1e-10)
    assert np.allclose(np.sort(_evals(GOLDEN["small_n"]).real), [1, 2], atol=1e-10)
    assert np.allclose(np.sort(_evals(GOLDEN["clustered"]).real), [1.0, 1.001, 1.002], atol=1e-8)
    for e in _evals(GOLDEN["repeated"]):
        assert min(abs(e - x) for x in (1, 2, 3, 4)) < 1e-8
    sv = GOLDEN["start_vector"]
    assert sv["iters"] == 1 and abs(_evals(sv)[0] - 0.1) < 1e-12
    interior = np.sort(_evals(GOLDEN["diagonal_interior_sigma"]).real)
    d = np.abs(0.1 * np.arange(1, 51) - 5.05)
    allowed = 0.1 * np.arange(1, 51)[d <= np.sort(d)[2] + 1e-12]
    assert all(np.min(np.abs(allowed - x)) < 1e-10 for x in interior)


@pytest.mark.parametrize("name", ["config1_n2000", "config5_like_n4096", "config4_like_n20000"])
def test_fixture_eigenvalues_are_pencil_eigenvalues(name):
    """every converged value of the banded+tau fixtures is an eigenvalue of
    the sparse pencil found independently by ARPACK on the shift-and-invert
    operator (A - sigma B)^-1 B (the oracle's LAPACK band LU, itself pinned
    against scipy spsolve in test_oracle_math): lambda = sigma + 1/mu, to
    1e-9 relative"""
    from scipy.sparse.linalg import LinearOperator, eigs
    from bbjd_amd.pencil import BandedTauPencil
    g = GOLDEN[name]
    p = BandedTauPencil.generate(**{k: g["pencil"][k] for k in ("n", "bs", "hw", "ntau", "seed")})
    A, B = p.to_scipy()
    s = complex(*g["sigma"])
    lu = orc.CpuPencilOps(p, s, 1)
    op = LinearOperator(A.shape, matvec=lambda x: lu.apply("orc_precond", B @ x), dtype=np.complex128)
    mu = eigs(op, k=g["nev"] + 6, which="LM", return_eigenvectors=False, tol=1e-14,
              ncv=max(4 * g["nev"] + 12, 40))
    w = s + 1.0 / mu
    for e in _evals(g):
        assert np.min(np.abs(w - e)) / abs(e) <= 1e-9


def test_fixture_random_sparse_matches_arpack():
    """reference tests/test_bbjd.py:235-278: eigs(sigma) agreement to 1e-8"""
    from scipy.sparse.linalg import eigs
    from tests.helpers import make_random_sparse
    A, _ = make_random_sparse(200)
    ref = eigs(A, k=6, sigma=5.0, which="LM", return_eigenvectors=False)
    for name in ("random_6", "random_gmres_inner20", "random_gmres_inner0"):
        for e in _evals(GOLDEN[name]):
            assert np.min(np.abs(ref - e)) < 1e-8


@pytest.mark.parametrize("which", ["oracle", "ref"])
def test_config3_block_jacobi_stalls_in_the_reference_algorithm(which):
    """Config 3 as specified -- GMRES(50) at inner_tol 1e-3 with the
    block-Jacobi preconditioner (inverses of the 8 x 8 diagonal blocks of
    A - sigma B) -- at the reference fixture's shape (n = 8000, block size 8,
    block-pentadiagonal, 8 tau rows, sigma = 1+2i, k = 6, maxsub 48, minsub
    20): the reference algorithm converges no pair in 100 outer iterations,
    its residual stays orders of magnitude above tol, while the same solve
    with the exact LU at sigma + 0.5 (the fixture config3_reduced_gmres)
    converges all 6.  This is why the GPU tests of config 3 use the
    approximate exact LU and test the block-Jacobi complex64 inner solve on
    its own."""
    if which == "ref" and not orc.have_ref():
        pytest.skip("reference build not present")
    from bbjd_amd.pencil import BandedTauPencil
    orc.load_oracle().orc_set_threads(4)
    p = BandedTauPencil.generate(8000, 8, 2, 8, seed=3)
    sigma = 1 + 2j
    ops = orc.CpuPencilOps(p, sigma, 2)
    r = orc.solve(which, p.n, 6, sigma, tol=1e-10, maxsub=48, minsub=20, maxiter=100, inner_maxiter=50,
                  inner_tol=1e-3, ops=ops, trace_cap=100)
    assert r["status"] == 1 and r["nconv"] == 0 and r["iters"] == 100
    if which == "oracle":
        res = r["trace"][:, 5]
        assert res.min() > 1e-4
    g = GOLDEN["config3_reduced_gmres"]
    assert g["nconv"] == 6 and g["status"] == 0
