"""Shared helpers for the tests (problem builders and checks)."""
import numpy as np
import scipy.sparse as sp


def make_diagonal(n, spacing=0.1):
    """eigenvalues spacing*[1..n] (reference tests/test_bbjd.py:109-112)"""
    d = np.arange(1, n + 1, dtype=np.complex128) * spacing
    return sp.diags(d, format='csr'), sp.eye(n, format='csr', dtype=np.complex128)


def make_random_sparse(n, sigma_shift=5.0, density=0.05, seed=42):
    """random non-Hermitian sparse (reference tests/test_bbjd.py:115-126)"""
    rng = np.random.RandomState(seed)
    d = min(density, 50.0 / n) if n > 1000 else density
    A = sp.random(n, n, density=d, format='csr', dtype=np.complex128, random_state=rng)
    A = A + sigma_shift * sp.eye(n, dtype=np.complex128)
    A = A.tocsr()
    A.data += 0.1j * rng.randn(A.nnz)
    B = sp.eye(n, format='csr', dtype=np.complex128)
    return A, B


def residual(A, B, lam, x):
    """||Ax - lam Bx|| / ||x|| (reference tests/test_bbjd.py:96-98)"""
    return np.linalg.norm(A @ x - lam * (B @ x)) / np.linalg.norm(x)


def match_error(ours, ref):
    """max over ours of the distance to the nearest ref value"""
    return max(min(abs(e - r) for r in ref) for e in ours)


def rel_match(ours, ref):
    """max relative distance of each of ours to its nearest ref value"""
    return max(min(abs(e - r) for r in ref) / max(abs(e), 1e-300) for e in ours)


def pencil_residual(A, B, lam, x, normA, normB):
    """north-star residual ||Ax - lam Bx|| / ((||A|| + |lam| ||B||) ||x||)"""
    return np.linalg.norm(A @ x - lam * (B @ x)) / ((normA + abs(lam) * normB) * np.linalg.norm(x))


def norm2_est(M, iters=20, seed=0):
    """lower bound of ||M||_2 by power iteration on M^H M"""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(M.shape[1]) + 1j * rng.standard_normal(M.shape[1])
    x /= np.linalg.norm(x)
    s = 0.0
    for _ in range(iters):
        y = M @ x
        z = M.conj().T @ y
        s = np.sqrt(np.linalg.norm(z))
        nz = np.linalg.norm(z)
        if nz == 0:
            return 0.0
        x = z / nz
    return float(np.linalg.norm(M @ x))
