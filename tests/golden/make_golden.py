"""Writes tests/golden/reference_runs.json: outputs of the reference solver
compiled from its own sources (oracle/_ref/libbbjd_ref.so, built by
oracle/Makefile from the reference's src/jd_core.c) on

  * the problems of the reference's own test suite (tests/test_bbjd.py:
    diagonal spectra, complex shift, clustered and repeated eigenvalues,
    tiny n, start vector, no preconditioner, random sparse vs ARPACK, the
    GMRES(20) approximate-preconditioner case), with scipy's splu as the
    preconditioner exactly as that suite does without FPLUS (:62-73);
  * the seeded banded+tau pencils of the benchmark configurations at full
    (config 1) or reduced size, with the CPU exact banded LU.

Recorded per case: converged eigenvalues in convergence order, residuals,
nconv, iters, status.  Run in the build container (where the reference
sources exist):  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import splu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bbjd_amd.pencil import BandedTauPencil  # noqa: E402
from oracle import orc  # noqa: E402
from tests.helpers import make_diagonal, make_random_sparse  # noqa: E402


def scipy_case(A, B, pc_matrix, **kw):
    A = sp.csc_matrix(A)
    B = sp.csc_matrix(B)
    lu = splu(sp.csc_matrix(pc_matrix)) if pc_matrix is not None else None
    return dict(matvec_A=lambda x: A @ x, matvec_B=lambda x: B @ x,
                precond=(lambda x: lu.solve(x)) if lu is not None else None, _mats=(A, B, lu), **kw)


def cases():
    """(name, builder) -> dict of orc.solve keyword arguments"""
    out = []

    def diag(name, n, k, sigma, pc_sigma=None, **kw):
        def b():
            A, B = make_diagonal(n)
            pcm = (A - (sigma if pc_sigma is None else pc_sigma) * B) if pc_sigma != "none" else None
            return dict(n=n, nev=k, sigma=sigma, **scipy_case(A, B, pcm, **kw))
        out.append((name, b))

    diag("diagonal_5_nearest_sigma0", 50, 5, 0.0, maxiter=200)              # test_bbjd.py:162
    diag("diagonal_interior_sigma", 50, 3, 5.05)                            # :176
    diag("request_subset", 20, 5, 0.0, maxiter=200)                         # :349
    diag("return_info_false", 20, 2, 0.0)                                   # :393
    diag("no_preconditioner", 20, 2, 0.0, pc_sigma="none", tol=1e-8, maxiter=500)  # :321

    def small_n():
        A = sp.diags([1.0, 2.0, 3.0, 4.0, 5.0]).astype(np.complex128)
        B = sp.eye(5, dtype=np.complex128)
        return dict(n=5, nev=2, sigma=0.0, **scipy_case(A, B, A))
    out.append(("small_n", small_n))                                        # :335

    def complex_sigma():
        n = 30
        rng = np.random.RandomState(7)
        d = np.arange(1, n + 1, dtype=np.complex128) + 0.1j * rng.randn(n)
        A = sp.diags(d)
        B = sp.eye(n, dtype=np.complex128)
        s = 15.0 + 0.1j
        return dict(n=n, nev=3, sigma=s, **scipy_case(A, B, A - s * B))
    out.append(("complex_sigma", complex_sigma))                            # :362

    def clustered():
        d = np.array([1.0, 1.001, 1.002, 2.0, 3.0, 4.0, 5.0], dtype=np.complex128)
        A = sp.diags(d)
        B = sp.eye(7, dtype=np.complex128)
        s = 0.9995
        return dict(n=7, nev=3, sigma=s, maxiter=500, maxsub=30, minsub=10,
                    **scipy_case(A, B, A - s * B))
    out.append(("clustered", clustered))                                    # :509

    def repeated():
        d = np.array([1.0, 2.0, 2.0, 3.0, 4.0], dtype=np.complex128)
        A = sp.diags(d)
        B = sp.eye(5, dtype=np.complex128)
        return dict(n=5, nev=3, sigma=2.0, maxiter=300, **scipy_case(A, B, A))
    out.append(("repeated", repeated))                                      # :527

    def start_vector():
        A, B = make_diagonal(20)
        v0 = np.zeros(20, np.complex128)
        v0[0] = 1.0
        return dict(n=20, nev=1, sigma=0.0, v0=v0, **scipy_case(A, B, A))
    out.append(("start_vector", start_vector))                              # :444

    def random_sparse(k, inner=0, pc_shift=0.0, maxsub=150, minsub=60, maxiter=500):
        def b():
            A, B = make_random_sparse(200)
            s = 5.0 + 0j
            kw = dict(maxiter=maxiter, maxsub=maxsub, minsub=minsub)
            if inner:
                kw = dict(maxiter=300, inner_maxiter=inner, inner_tol=1e-2)
            return dict(n=200, nev=k, sigma=s, **kw, **scipy_case(A, B, A - (s + pc_shift) * B))
        return b
    out.append(("random_6", random_sparse(6)))                              # :235
    out.append(("random_gmres_inner0", random_sparse(2, 0, 0.5, maxiter=300)))     # :459
    out.append(("random_gmres_inner20", random_sparse(2, 20, 0.5)))         # :459
    out.append(("random_restart_maxsub30", random_sparse(4, 0, 0.0, 30, 15, 2000)))  # :601

    def banded(name, n, bs, hw, ntau, seed, sigma, k, maxsub, minsub, inner=0, pc_shift=0.0, inner_tol=1e-3,
               maxiter=1000):
        def b():
            p = BandedTauPencil.generate(n, bs, hw, ntau, seed=seed)
            ops = orc.CpuPencilOps(p, sigma + pc_shift, 1)
            return dict(n=n, nev=k, sigma=sigma, maxsub=maxsub, minsub=minsub, maxiter=maxiter,
                        inner_maxiter=inner, inner_tol=inner_tol, ops=ops, _keep=(p, ops),
                        _pencil=dict(n=n, bs=bs, hw=hw, ntau=ntau, seed=seed, pc_shift=pc_shift))
        out.append((name, b))

    banded("config1_n2000", 2000, 4, 1, 2, 1, 0.5 + 0.5j, 8, 40, 20)
    banded("config2_reduced_n16000", 16000, 8, 2, 16, 2, 3.0 + 0j, 12, 60, 24)
    banded("config3_reduced_gmres", 8000, 8, 2, 8, 3, 1 + 2j, 6, 48, 20, inner=50, pc_shift=0.5)
    banded("config4_like_n20000", 20000, 8, 1, 4, 4, -1.5 + 2.0j, 8, 32, 16)
    banded("config5_like_n4096", 4096, 4, 1, 2, 5, 0.3 - 0.6j, 8, 32, 16)
    return out


def run_case(builder, which):
    kw = builder()
    kw.pop("_keep", None)
    kw.pop("_mats", None)
    pen = kw.pop("_pencil", None)
    r = orc.solve(which, **kw)
    rec = {k: kw[k] for k in ("n", "nev", "maxsub", "minsub", "maxiter", "inner_maxiter", "tol",
                              "inner_tol") if k in kw}
    rec["sigma"] = [complex(kw["sigma"]).real, complex(kw["sigma"]).imag]
    rec.update(nconv=r["nconv"], iters=r["iters"], status=r["status"],
               evals=[[float(e.real), float(e.imag)] for e in r["evals"]],
               residuals=[float(x) for x in r["resid"]])
    if pen:
        rec["pencil"] = pen
    return rec


def main():
    if not orc.have_ref():
        sys.exit("oracle/_ref/libbbjd_ref.so is not built (make -C oracle with the reference sources)")
    orc.load_oracle().orc_set_threads(1)
    data = {"generator": "tests/golden/make_golden.py", "solver": "reference src/jd_core.c via oracle/_ref",
            "blas": "scipy bundled OpenBLAS 0.3.28, 1 thread", "cases": {}}
    for name, b in cases():
        data["cases"][name] = run_case(b, "ref")
        c = data["cases"][name]
        print(f"{name}: nconv {c['nconv']} iters {c['iters']} status {c['status']}", flush=True)
    with open(os.path.join(HERE, "reference_runs.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
