"""The projected step on the device (bbjd_qz_device_*: one wavefront per
problem) against LAPACK and against the reference's selection and target
rules restated by brute force (reference src/jd_core.c:319-462), and the
engine with qz_mode='device' against the reference's own runs.

Bars: eigenvalues within 1e-12 relative of scipy.linalg.eigvals(MA, MB)
(both backward stable; the pencils are well conditioned); Z unitary and
Z T Z^H = C^-1 MB to 1e-12 relative; T upper triangular to 1e-14; the Ritz
coefficients satisfy MA x = theta MB x to 1e-10 relative; selection, order
and target exactly as the brute-force rule; end to end the golden bars of
test_gpu_configs."""
import ctypes as C
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl

from bbjd_amd import BandedTauPencil, DeviceOperator, _lib
from bbjd_amd._bbjd_c import jdqz_c
from tests.helpers import rel_match

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "reference_runs.json")))["cases"]
OK, AMBIGUOUS, NOT_TRIANGULAR, ACCURACY, NO_CONVERGENCE, UNSUPPORTED = range(6)
TARGETS = {"nearest": 0, "LM": 1, "SM": 2, "LR": 3, "SR": 4, "LI": 5, "SI": 6}


def device_step(MA, MB, sigma, conv=(), nkeep=1, target=0, tol=1e-10, nu0=None, mu0=None):
    lib = _lib.load()
    f = lib.bbjd_qz_device_host
    f.restype = C.c_int
    Z = _lib.zcomplex
    f.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_int64, Z, Z, Z, C.c_int, C.c_int64, C.c_void_p, C.c_int64,
                  C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    k = MA.shape[0]
    if nu0 is None:
        nu0 = 1 / np.sqrt(1 + abs(sigma) ** 2)
        mu0 = -sigma * nu0
    MAf, MBf = np.asfortranarray(MA, np.complex128), np.asfortranarray(MB, np.complex128)
    cv = np.ascontiguousarray(np.asarray(conv, np.complex128)) if len(conv) else np.zeros(1, np.complex128)
    al, be, x = (np.zeros(k, np.complex128) for _ in range(3))
    Zm, T = np.zeros((k, k), np.complex128, order="F"), np.zeros((k, k), np.complex128, order="F")
    info = np.zeros(4, np.int32)
    rc = f(k, MAf.ctypes.data, MBf.ctypes.data, k, _lib.zc(nu0), _lib.zc(mu0), _lib.zc(sigma), target, nkeep,
           cv.ctypes.data, len(conv), tol, al.ctypes.data, be.ctypes.data, x.ctypes.data, Zm.ctypes.data,
           T.ctypes.data, info.ctypes.data)
    assert rc == 0
    return dict(al=al, be=be, x=x, Z=Zm, T=T, status=int(info[0]), p=int(info[1]), nsel=int(info[2]),
                its=int(info[3]), nu0=nu0, mu0=mu0)


def projected(k, seed, sigma):
    """a projected pencil as the engine forms it: V orthonormal, W the
    orthonormal factor of (nu0 A + mu0 B) V, MA = W^H A V, MB = W^H B V (so
    C = nu0 MA + mu0 MB is upper triangular)"""
    rng = np.random.default_rng(seed)
    n = 3 * k + 4
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = np.diag(1.0 + 0.1 * rng.standard_normal(n))
    V, _ = np.linalg.qr(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))
    nu0 = 1 / np.sqrt(1 + abs(sigma) ** 2)
    W, _ = np.linalg.qr((nu0 * A - sigma * nu0 * B) @ V)
    return W.conj().T @ A @ V, W.conj().T @ B @ V


def key(t, lam, sigma):
    return {0: abs(lam - sigma), 1: -abs(lam), 2: abs(lam), 3: -lam.real, 4: lam.real, 5: -lam.imag,
            6: lam.imag}[t]


def brute(theta, conv, nkeep, target, sigma, tol=1e-10):
    """the reference's rule: selected = matching a converged value within
    100 tol, then the nkeep best by key; target = best unconverged value"""
    conv = np.asarray(conv, np.complex128)
    isc = np.array([bool(len(conv)) and np.min(np.abs(conv - t)) < 100 * tol for t in theta])
    sel = set(np.flatnonzero(isc))
    for _ in range(nkeep):
        rest = [i for i in range(len(theta)) if i not in sel]
        if not rest:
            break
        sel.add(min(rest, key=lambda i: key(target, theta[i], sigma)))
    cand = [i for i in range(len(theta)) if not isc[i]]
    best = min(cand, key=lambda i: key(target, theta[i], sigma)) if cand else None
    return sel, best


@pytest.mark.parametrize("k", [1, 2, 3, 8, 17, 32, 48, 64])
def test_device_step_matches_lapack(k):
    sigma = 0.3 - 0.2j
    MA, MB = projected(k, 100 + k, sigma)
    ref = sl.eigvals(MA, MB)
    nu0 = 1 / np.sqrt(1 + abs(sigma) ** 2)
    conv = list(ref[:2]) if k > 4 else []
    nkeep = max(1, k // 3)
    out = device_step(MA, MB, sigma, conv, nkeep)
    assert out["status"] in (OK, AMBIGUOUS), out["status"]
    th = out["al"] / out["be"]
    for t in th:
        assert np.min(np.abs(ref - t)) <= 1e-12 * abs(t), (t, out["its"])
    Zm, T = out["Z"], out["T"]
    M = np.linalg.solve(np.triu(nu0 * MA - sigma * nu0 * MB), MB)
    assert np.abs(Zm.conj().T @ Zm - np.eye(k)).max() <= 1e-13
    assert np.abs(np.tril(T, -1)).max() <= 1e-14 * np.abs(T).max()
    assert np.linalg.norm(Zm @ T @ Zm.conj().T - M) <= 1e-12 * np.linalg.norm(M)
    # selection at the top in order of appearance, the target by the rule
    sel, best = brute(th, conv, nkeep, 0, sigma)
    assert out["nsel"] == len(sel)
    assert sel == set(range(len(sel)))                  # the selected values lead
    if out["status"] == OK:
        assert out["p"] == best
        x = out["x"]
        r = MA @ x - th[out["p"]] * (MB @ x)
        assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(MA) * np.linalg.norm(x)


@pytest.mark.parametrize("target", sorted(TARGETS))
def test_device_selection_and_target_rules_by_brute_force(target):
    """diagonal pencils with known values: the selected set, their order
    (as they appear), and the target for every target key, against the
    brute-force rule; a tie for the target and an all-converged spectrum are
    flagged (the host decides those)"""
    rng = np.random.default_rng(TARGETS[target])
    t = TARGETS[target]
    for trial in range(6):
        k = int(rng.integers(3, 40))
        theta = np.round(rng.standard_normal(k) + 1j * rng.standard_normal(k), 3)
        theta = np.unique(theta)
        k = len(theta)
        sigma = 0.1 + 0.2j
        conv = list(rng.choice(theta, size=min(3, k - 1), replace=False)) if trial % 2 else []
        nkeep = int(rng.integers(1, k))
        out = device_step(np.diag(theta), np.eye(k), sigma, conv, nkeep, target=t)
        th = out["al"] / out["be"]
        sel, best = brute(theta, conv, nkeep, t, sigma)
        # the leading nsel values are the selected ones, in their original order
        assert out["nsel"] == len(sel)
        assert np.allclose(th[: len(sel)], theta[sorted(sel)], atol=1e-12)
        keys = sorted(key(t, theta[i], sigma) for i in range(k)
                      if not (conv and np.min(np.abs(np.asarray(conv) - theta[i])) < 1e-8))
        tie = len(keys) > 1 and keys[1] - keys[0] <= 1e-10 * max(abs(keys[0]), abs(keys[1]), 1e-14)
        if tie:
            assert out["status"] == AMBIGUOUS
        else:
            assert out["status"] == OK
            assert abs(th[out["p"]] - theta[best]) <= 1e-12 * max(1, abs(theta[best]))
    # every value converged: nothing to target
    theta = np.array([1.0, 2.0, 3.0]) + 0j
    assert device_step(np.diag(theta), np.eye(3), 0.5, list(theta), 1)["status"] == AMBIGUOUS
    # an exact tie: 1 and 3 around sigma = 2
    theta = np.array([1.0, 3.0, 5.0]) + 0j
    assert device_step(np.diag(theta), np.eye(3), 2.0, [], 1)["status"] == AMBIGUOUS


def test_device_step_flags_a_non_triangular_test_space():
    rng = np.random.default_rng(3)
    MA = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    MB = np.eye(6) + 0j
    assert device_step(MA, MB, 0.5)["status"] == NOT_TRIANGULAR


@pytest.mark.parametrize("name", ["config1_n2000", "config3_reduced_gmres", "config4_like_n20000",
                                  "config5_like_n4096", "config2_reduced_n16000"])
def test_engine_with_device_projected_step_matches_reference(name):
    """qz_mode='device': the projected step of every iteration with k <= 64
    on the GPU (restarts included: their rotation is the device's Schur
    vectors); the reference's status, nconv and eigenvalues (1e-8)"""
    g = GOLDEN[name]
    pen = g["pencil"]
    p = BandedTauPencil.generate(pen["n"], pen["bs"], pen["hw"], pen["ntau"], seed=pen["seed"])
    sigma = complex(*g["sigma"])
    dev = DeviceOperator(p, sigma + pen["pc_shift"], pc="exact")
    out = jdqz_c(p.n, g["nev"], sigma, 0, None, None, True, tol=1e-10, maxiter=g["maxiter"], maxsub=g["maxsub"],
                 minsub=g["minsub"], inner_maxiter=g["inner_maxiter"], inner_tol=g["inner_tol"], device_op=dev,
                 qz_mode="device", return_stats=True)
    ev, X, res, nconv, iters, status, st = out
    print(f"[device qz] {name}: nconv {nconv} iters {iters} (ref {g['iters']}); device steps "
          f"{st['qz_device_steps']}, to the host {st['qz_device_fallbacks']} ({st['qz_device_accuracy']} by the "
          f"accuracy bound)", flush=True)
    assert status == g["status"] and nconv == g["nconv"]
    ge = np.array([complex(a, b) for a, b in g["evals"]])
    assert rel_match(ev, ge) <= 1e-8 and rel_match(ge, ev) <= 1e-8
    assert np.all(res < 1e-10)
    if g["maxsub"] <= 64:
        # every iteration went to the device; the ones it flagged (a theta
        # whose error bound exceeds 10 tol, a tie) were decided on the host
        assert st["qz_device_steps"] == iters
        assert st["qz_device_fallbacks"] < st["qz_device_steps"]
    A, B = p.to_scipy()
    for j in range(nconv):
        assert np.linalg.norm(A @ X[:, j] - ev[j] * (B @ X[:, j])) <= 1e-10 * (1 + abs(ev[j])) * 10


def test_batched_launch_equals_one_problem_launches():
    """many problems of different sizes, targets and flags in one launch:
    each wavefront's result bit-identical to the same problem launched alone,
    and the restart rotation the leading keep columns of its Z"""
    from bbjd_amd.qz_device import step_batch
    rng = np.random.default_rng(5)
    probs = []
    for i, k in enumerate([1, 5, 64, 17, 33, 8, 64, 2, 48, 40, 12, 64]):
        sigma = complex(rng.standard_normal(), rng.standard_normal())
        MA, MB = projected(k, 300 + i, sigma)
        conv = list(sl.eigvals(MA, MB)[:1]) if k > 6 and i % 2 else []
        probs.append(dict(MA=MA, MB=MB, sigma=sigma, conv=conv, nkeep=max(1, k // 3), target=i % 7,
                          keep=max(1, k // 2)))
    probs.insert(3, dict(MA=rng.standard_normal((6, 6)) + 0j, MB=np.eye(6) + 0j, sigma=0.5))   # not triangular
    batch = step_batch(probs, full=True)
    assert batch[3]["status"] == NOT_TRIANGULAR
    for q, got in zip(probs, batch):
        one = device_step(q["MA"], q["MB"], q["sigma"], q.get("conv", ()), q.get("nkeep", 1), q.get("target", 0))
        assert got["status"] == one["status"] and got["p"] == one["p"] and got["nsel"] == one["nsel"]
        if got["status"] != NOT_TRIANGULAR:
            for key_b, key_o in (("alpha", "al"), ("beta", "be"), ("x", "x"), ("Z", "Z"), ("T", "T")):
                assert np.array_equal(got[key_b], one[key_o]), key_b
            assert np.array_equal(got["zr"], got["Z"][:, : q["keep"]])
