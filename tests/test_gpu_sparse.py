"""General sparse operators on the GPU (include/bbjd_sparse.h): the CSR
matvec against scipy, the sparse-LU apply against SuperLU's own solve, and
the reference's own scipy cases (tests/golden/reference_runs.json, written
by the reference solver with scipy matvecs and splu) solved with every
operator on the device.

Bars: matvec and preconditioner apply within 1e-13 / 1e-12 relative (the
same products and substitutions, summed in a different order); solves as
in test_gpu_configs: same status and nconv, eigenvalues within 1e-8 of the
reference's in both directions, every pair's residual from scipy on the
host within tol (1 + |lambda|)."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import splu

from bbjd_amd import SparseDeviceOperator, jdqz
from tests.golden.make_golden import cases
from tests.helpers import rel_match

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "reference_runs.json")))["cases"]
CASES = dict(cases())
SCIPY_CASES = sorted(k for k, v in GOLDEN.items() if "pencil" not in v)


def _random(n, density, seed, shift=3.0):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=seed, format="csr").astype(np.complex128)
    M.data += 0.3j * rng.standard_normal(M.nnz)
    return (M + shift * sp.eye(n)).tocsr()


def _matrices():
    rng = np.random.default_rng(7)
    out = {"random_1e3": _random(1000, 0.01, 1), "diagonal_5e3": sp.diags(rng.standard_normal(5000) + 2j),
           "one": sp.csr_matrix(np.array([[2.0 - 1j]])), "wide_rows": _random(600, 0.4, 2)}
    # empty rows and one dense row
    M = _random(5000, 0.0008, 3).tolil()
    M[17, :] = 0
    M[4000, :] = rng.standard_normal(5000) + 1j * rng.standard_normal(5000)
    out["empty_and_dense_rows"] = M.tocsr()
    return out


@pytest.mark.parametrize("name", sorted(_matrices()))
def test_csr_matvec_matches_scipy(name):
    M = _matrices()[name]
    n = M.shape[0]
    Bm = _random(n, min(1.0, 5.0 / n), 9) if n > 1 else sp.eye(1, dtype=np.complex128)
    op = SparseDeviceOperator(M, Bm)
    opI = SparseDeviceOperator(M)
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for got, ref in ((op.matvec(x), M @ x), (op.matvec_B(x), Bm @ x), (opI.matvec_B(x), x)):
        assert np.linalg.norm(got - ref) <= 1e-13 * max(np.linalg.norm(ref), 1e-300)


@pytest.mark.parametrize("case", ["random_2e3", "tridiagonal_5e3", "diagonal_1e4", "dense_row_3e3"])
def test_sparse_lu_apply_matches_superlu(case):
    """L and U of splu(M) on the device: the level-set solves (a chain of
    about 5000 one-row levels walked by one workgroup for the tridiagonal case;
    one 10000-row level as a grid for the diagonal; mixed for the others)
    reproduce SuperLU's solve"""
    rng = np.random.default_rng(11)
    if case == "random_2e3":
        M = _random(2000, 0.003, 4)
    elif case == "tridiagonal_5e3":
        n = 5000
        M = sp.diags([rng.standard_normal(n - 1) + 0j, 4 + rng.standard_normal(n) + 1j, rng.standard_normal(n - 1) + 0j],
                     [-1, 0, 1], format="csr")
    elif case == "diagonal_1e4":
        M = sp.diags(rng.standard_normal(10000) + 3 + 1j, format="csr")
    else:
        M = _random(3000, 0.001, 5).tolil()
        M[2999, :] = rng.standard_normal(3000) + 1j * rng.standard_normal(3000)
        M = M.tocsr()
    n = M.shape[0]
    lu = splu(M.tocsc())
    op = SparseDeviceOperator(M, None, lu)
    lev_l, lev_u, launches = op.plan()
    print(f"[sparse lu] {case}: L levels {lev_l}, U levels {lev_u}, launches per apply {launches}", flush=True)
    for seed in (1, 2):
        r = np.random.default_rng(seed)
        b = r.standard_normal(n) + 1j * r.standard_normal(n)
        ref = lu.solve(b)
        assert np.linalg.norm(op.solve(b) - ref) <= 1e-12 * np.linalg.norm(ref)
    if case == "tridiagonal_5e3":
        assert lev_l > 4900 and launches == 4        # a chain: one walker launch per factor, two gathers
    if case == "diagonal_1e4":
        assert lev_l == lev_u == 1


@pytest.mark.parametrize("name", SCIPY_CASES)
def test_golden_scipy_case_on_device_sparse_operators(name):
    """the reference's scipy-matrix cases with A, B and the same splu
    preconditioner as device operators: the reference's status, nconv and
    eigenvalues"""
    g = GOLDEN[name]
    kw = CASES[name]()
    A, B, lu = kw["_mats"]
    op = SparseDeviceOperator(A, B, lu)
    tol = kw.get("tol", 1e-10)
    ev, X, info = jdqz(op, k=kw["nev"], sigma=kw["sigma"], tol=tol, maxiter=kw.get("maxiter", 200),
                       maxsub=kw.get("maxsub", 0), minsub=kw.get("minsub", 0),
                       inner_maxiter=kw.get("inner_maxiter", 0), inner_tol=kw.get("inner_tol", 1e-2),
                       preconditioner=True if lu is not None else None, v0=kw.get("v0"), return_info=True)
    print(f"[sparse golden] {name}: nconv {info['nconv']} (ref {g['nconv']}), iters {info['iters']} "
          f"(ref {g['iters']})", flush=True)
    assert info["status"] == g["status"] and info["nconv"] == g["nconv"]
    ge = np.array([complex(a, b) for a, b in g["evals"]])
    assert rel_match(ev, ge) <= 1e-8 and rel_match(ge, ev) <= 1e-8
    for j in range(info["nconv"]):
        x = X[:, j]
        r = np.linalg.norm(A @ x - ev[j] * (B @ x)) / np.linalg.norm(x)
        assert r <= 2 * tol * (1 + abs(ev[j]))


def test_jdqz_device_flag_moves_scipy_operands_to_the_gpu():
    """the reference's call with scipy matrices and a SuperLU, plus
    device=True: the same eigenvalues as the host-callback path"""
    from tests.helpers import make_random_sparse
    A, B = make_random_sparse(200)
    lu = splu((A - 5.0 * B).tocsc())
    kw = dict(k=4, sigma=5.0, tol=1e-10, maxsub=30, minsub=15, maxiter=2000, preconditioner=lu)
    e_dev = jdqz(A, B, device=True, **kw)[0]
    e_host = jdqz(A, B, **kw)[0]
    assert len(e_dev) == len(e_host) == 4
    assert rel_match(e_dev, e_host) <= 1e-8 and rel_match(e_host, e_dev) <= 1e-8


@pytest.mark.parametrize("n", [2000, 20000, 100000])
def test_pinned_dense_rows_apply(n):
    """a banded pencil with 4 dense rows given as scipy matrices and the
    matrix A - sigma B to factor: the dense rows are pinned out of the sparse
    LU and enter through the Schur complement on the device; the apply
    solves M y = b to 1e-12 (residual from scipy's matvec), agrees with the
    host form of the same factors and with the banded operator's exact
    block-cyclic-reduction solve (1e-10)"""
    from bbjd_amd import BandedTauPencil, DeviceOperator
    from bbjd_amd.sparse import bordered_solve, pinned_factors
    p = BandedTauPencil.generate(n, 8, 1, 4, seed=3)
    A, B = p.to_scipy()
    sigma = 0.5 + 0.5j
    M = (A - sigma * B).tocsr()
    op = SparseDeviceOperator(A, B, M)
    assert op.pinned_rows.tolist() == list(range(p.n1, p.n))
    lev_l, lev_u, launches = op.plan()
    print(f"[pinned] n={n}: L levels {lev_l}, U levels {lev_u}, launches per apply {launches}, "
          f"device bytes {op.device_bytes}", flush=True)
    rng = np.random.default_rng(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = op.solve(b)
    assert np.linalg.norm(M @ y - b) <= 1e-12 * np.linalg.norm(b) * max(1.0, np.abs(M).max())
    if n <= 20000:
        yh = bordered_solve(pinned_factors(M, op.pinned_rows), b)
        assert np.linalg.norm(y - yh) <= 1e-12 * np.linalg.norm(yh)
    dev = DeviceOperator(p, sigma, pc="exact")
    try:
        yb = dev.solve(b)
    finally:
        dev.free()
    assert np.linalg.norm(y - yb) <= 1e-10 * np.linalg.norm(yb)


@pytest.mark.parametrize("name", ["config1_n2000", "config4_like_n20000"])
def test_golden_banded_case_on_pinned_sparse_operators(name):
    """the reference's banded runs (dense tau rows) with A and B as scipy
    matrices on the device and A - sigma B factored with its dense rows
    pinned: the reference's status, nconv and eigenvalues (1e-8)"""
    from bbjd_amd import BandedTauPencil
    g = GOLDEN[name]
    pen = g["pencil"]
    p = BandedTauPencil.generate(pen["n"], pen["bs"], pen["hw"], pen["ntau"], seed=pen["seed"])
    A, B = p.to_scipy()
    sigma = complex(*g["sigma"])
    op = SparseDeviceOperator(A, B, (A - (sigma + pen["pc_shift"]) * B).tocsr())
    assert op.pinned_rows.size == pen["ntau"]
    ev, X, info = jdqz(op, k=g["nev"], sigma=sigma, tol=1e-10, maxiter=g["maxiter"], maxsub=g["maxsub"],
                       minsub=g["minsub"], preconditioner=True, return_info=True)
    print(f"[pinned golden] {name}: nconv {info['nconv']} iters {info['iters']} (ref {g['iters']})", flush=True)
    assert info["status"] == g["status"] and info["nconv"] == g["nconv"]
    ge = np.array([complex(a, b) for a, b in g["evals"]])
    assert rel_match(ev, ge) <= 1e-8 and rel_match(ge, ev) <= 1e-8
    for j in range(info["nconv"]):
        x = X[:, j]
        assert np.linalg.norm(A @ x - ev[j] * (B @ x)) / np.linalg.norm(x) <= 2e-10 * (1 + abs(ev[j]))
