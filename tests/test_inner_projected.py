"""The projected-preconditioner form of the inner solve (the engine's
BBJD_INNER_PROJECTED, complex64 on the GPU) checked in FP64 numpy at small
size against the dense solution of the correction equation the reference's
GMRES solves (reference src/jd_core.c:568-727):

    (I - Z Z^H)(tb A - ta B)(I - Q Q^H) t = -(I - Z Z^H) r,   Q^H t = 0.

With Yh = M^-1 Z, H = Q^H Yh and Pi = I - Yh H^-1 Q^H (Fokkema, Sleijpen and
van der Vorst, SIAM J. Sci. Comput. 20 (1998), sec. 3.3), GMRES runs on
Pi M^-1 (tb A - ta B) from Pi M^-1 (-r).  CPU only."""
import numpy as np


def _problem(n=72, K=4, seed=3):
    rng = np.random.default_rng(seed)
    cn = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    A = cn(n, n)
    B = np.eye(n) + 0.1 * cn(n, n)
    M = A - (0.3 + 0.2j) * B + 0.05 * rng.standard_normal((n, n))     # approximate A - sigma B
    Q = np.linalg.qr(cn(n, K))[0]
    Z = np.linalg.qr(cn(n, K))[0]
    ta, tb = 0.7 + 0.1j, 1.1 - 0.2j
    return A, B, M, Q, Z, ta, tb, cn(n)


def _gmres(op, b, m, tol=1e-12):
    """GMRES with twice-classical Gram-Schmidt, stopped when the least-squares
    residual falls below tol * ||b|| (as the engine stops at inner_tol);
    returns the iterate"""
    beta = np.linalg.norm(b)
    V = [b / beta]
    H = np.zeros((m + 1, m), complex)
    y = np.zeros(0, complex)
    for j in range(m):
        w = op(V[j])
        for _ in range(2):
            h = np.array([np.vdot(v, w) for v in V])
            H[:len(V), j] += h
            w = w - np.array(V).T @ h
        H[j + 1, j] = np.linalg.norm(w)
        e = np.zeros(j + 2, complex)
        e[0] = beta
        y, res = np.linalg.lstsq(H[:j + 2, :j + 1], e, rcond=None)[:2]
        if np.linalg.norm(H[:j + 2, :j + 1] @ y - e) <= tol * beta or H[j + 1, j] <= tol * beta:
            break
        V.append(w / H[j + 1, j])
    return np.array(V[:len(y)]).T @ y


def _dense_correction(A, B, Q, Z, ta, tb, r):
    n, K = Q.shape
    rng = np.random.default_rng(0)
    Qp = np.linalg.qr(np.hstack([Q, rng.standard_normal((n, n - K))]))[0][:, K:]
    Zp = np.linalg.qr(np.hstack([Z, rng.standard_normal((n, n - K))]))[0][:, K:]
    Ao = tb * A - ta * B
    return Qp @ np.linalg.solve(Zp.conj().T @ Ao @ Qp, -(Zp.conj().T @ r))


def test_projected_form_solves_the_correction_equation():
    A, B, M, Q, Z, ta, tb, r = _problem()
    n, K = Q.shape
    Ao, Mi = tb * A - ta * B, np.linalg.inv(M)
    Yh = Mi @ Z
    Hi = np.linalg.inv(Q.conj().T @ Yh)
    Pi = lambda y: y - Yh @ (Hi @ (Q.conj().T @ y))
    PR = np.eye(n) - Q @ Q.conj().T
    t = PR @ _gmres(lambda v: Pi(Mi @ (Ao @ v)), Pi(Mi @ -r), n - K)
    t_ex = _dense_correction(A, B, Q, Z, ta, tb, r)
    assert np.linalg.norm(t - t_ex) <= 1e-9 * np.linalg.norm(t_ex)
    # the reference form reaches the same solution
    PL = np.eye(n) - Z @ Z.conj().T
    t_ref = PR @ _gmres(lambda v: Mi @ (PL @ (Ao @ (PR @ v))), Mi @ (PL @ -r), n)
    assert np.linalg.norm(t_ref - t_ex) <= 1e-9 * np.linalg.norm(t_ex)


def test_projected_preconditioner_identities():
    """Pi M^-1 (I - Z Z^H) = Pi M^-1 (the left projection is implied), and
    Pi maps into Q-perp, so no right projection is needed on the Krylov
    vectors"""
    A, B, M, Q, Z, ta, tb, r = _problem(seed=5)
    n = Q.shape[0]
    Mi = np.linalg.inv(M)
    Yh = Mi @ Z
    Hi = np.linalg.inv(Q.conj().T @ Yh)
    Pi = np.eye(n) - Yh @ Hi @ Q.conj().T
    PL = np.eye(n) - Z @ Z.conj().T
    assert np.linalg.norm(Pi @ Mi @ PL - Pi @ Mi) <= 1e-10 * np.linalg.norm(Pi @ Mi)
    assert np.linalg.norm(Q.conj().T @ Pi) <= 1e-10 * np.linalg.norm(Pi)


def test_projected_form_converges_faster_with_an_approximate_preconditioner():
    """with M a perturbed A - sigma B and theta away from sigma, five steps of
    the projected form leave a smaller residual of the correction equation
    than five of the reference form (the effect that lets the served mixed
    form take GMRES(5))"""
    A, B, M, Q, Z, ta, tb, r = _problem(seed=7)
    n = Q.shape[0]
    Ao, Mi = tb * A - ta * B, np.linalg.inv(M)
    PL = np.eye(n) - Z @ Z.conj().T
    PR = np.eye(n) - Q @ Q.conj().T
    Yh = Mi @ Z
    Hi = np.linalg.inv(Q.conj().T @ Yh)
    Pi = lambda y: y - Yh @ (Hi @ (Q.conj().T @ y))
    res = lambda t: np.linalg.norm(PL @ (Ao @ t + r)) / np.linalg.norm(PL @ r)
    t_pp = PR @ _gmres(lambda v: Pi(Mi @ (Ao @ v)), Pi(Mi @ -r), 5)
    t_rf = PR @ _gmres(lambda v: Mi @ (PL @ (Ao @ (PR @ v))), Mi @ (PL @ -r), 5)
    assert res(t_pp) < res(t_rf)
