"""σ sweeps (bbjd_amd.sweep): the merge of eigenvalues found by several
requests on the CPU, and a served sweep with and without warm starts on the
GPU, each request checked against the oracle."""
import numpy as np
import pytest

from bbjd_amd.sweep import dedup


def test_dedup_merges_within_tolerance_and_keeps_smallest_residual():
    lam = [1.0 + 1.0j, 2.0, 1.0 + 1.0j + 5e-9, 3.0 - 1j, 1.0 + 1.0j - 3e-9, 2.0 + 1e-6]
    res = [3e-11, 1e-11, 1e-12, 2e-11, 5e-11, 4e-11]
    keep, groups = dedup(lam, res, rtol=1e-8)
    # 1+1i (three copies within 1e-8 |λ|), 2, 2 + 1e-6 (too far to merge), 3-1i
    assert len(keep) == 4
    assert list(keep) == [2, 1, 5, 3]                 # sorted by real part; 1e-12 represents the triple
    assert groups[0] == [0, 2, 4]
    assert all(len(g) == 1 for g in groups[1:])


def test_dedup_is_transitive_and_scale_aware():
    # a chain a ~ b ~ c with a, c farther apart than the tolerance merges
    lam = [10.0, 10.0 + 0.9e-7, 10.0 + 1.8e-7]
    keep, groups = dedup(lam, [1.0, 2.0, 3.0], rtol=1e-8)   # tolerance 1e-7 at |λ| = 10
    assert len(keep) == 1 and groups[0] == [0, 1, 2] and keep[0] == 0
    # below |λ| = 1 the tolerance is absolute
    keep, _ = dedup([1e-3, 1e-3 + 5e-9, 1e-3 + 2e-8], [1.0, 1.0, 1.0], rtol=1e-8)
    assert len(keep) == 2
    assert len(dedup([], [])[0]) == 0


@pytest.mark.gpu
def test_served_sweep_merges_overlapping_windows():
    """four σ points, two of them 1e-3 apart so that their windows of nev
    nearest eigenvalues overlap: the merged set has every pair's residual
    below tol (||A|| + |λ|), holds every request's eigenvalues, has no two
    within the merge tolerance, and is the same with warm starts from the
    first wave"""
    from bbjd_amd import BandedTauPencil
    from bbjd_amd.configs import row_norm_bound
    from bbjd_amd.sweep import sweep
    p = BandedTauPencil.generate(4000, 8, 1, 4, seed=17)
    sigmas = [0.5 + 0.5j, 0.501 + 0.5j, -2.0 + 1.0j, 2.5 - 1.5j]
    out = sweep(p, sigmas, nev=5, tol=1e-10, maxsub=20, minsub=10, workers=4)
    chained = sweep(p, sigmas, nev=5, tol=1e-10, maxsub=20, minsub=10, workers=4, chain=True)
    print(f"[sweep] merged {len(out['eigenvalues'])} duplicates {out['duplicates']}; chained merged "
          f"{len(chained['eigenvalues'])} iterations {[r['iters'] for r in out['requests']]} -> "
          f"{[r['iters'] for r in chained['requests']]}", flush=True)
    anorm = row_norm_bound(p)
    for o in (out, chained):
        assert all(r["status"] == 0 and r["nconv"] == 5 for r in o["requests"])
        lam = o["eigenvalues"]
        assert o["duplicates"] >= 3
        assert len(lam) + o["duplicates"] == 20
        assert np.all(o["residuals"] <= 1e-10 * (anorm + np.abs(lam)))
        for r in o["requests"]:
            for ev in r["eigenvalues"]:
                assert np.min(np.abs(lam - ev)) <= 1e-8 * max(1.0, abs(ev))
        d = np.abs(lam[:, None] - lam[None, :]) + np.eye(len(lam)) * 1e9
        assert d.min() > 1e-8
        A, B = p.to_scipy()
        X = o["eigenvectors"]
        for j in range(len(lam)):
            assert np.linalg.norm(A @ X[:, j] - lam[j] * (B @ X[:, j])) <= 1e-10 * (anorm + abs(lam[j]))
    assert len(out["eigenvalues"]) == len(chained["eigenvalues"])
    for ev in chained["eigenvalues"]:
        assert np.min(np.abs(out["eigenvalues"] - ev)) <= 1e-8 * max(1.0, abs(ev))
    # each request of the plain sweep against the oracle (CPU restatement of
    # the reference solver) at the same sigma, same settings and start vector:
    # the same eigenvalues within 1e-8 relative, in both directions
    from oracle import orc
    from tests.helpers import rel_match
    for s, r in zip(sigmas, out["requests"]):
        ref = orc.solve("oracle", p.n, 5, s, tol=1e-10, maxsub=20, minsub=10, maxiter=1000,
                        ops=orc.CpuPencilOps(p, s, 1))
        assert ref["nconv"] == r["nconv"] == 5
        assert rel_match(r["eigenvalues"], ref["evals"]) <= 1e-8
        assert rel_match(ref["evals"], r["eigenvalues"]) <= 1e-8
