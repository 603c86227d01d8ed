"""Run the reference's benchmark configurations at full size on one GPU.

    python -m bbjd_amd.configs CONFIG [--seed S] [--requests R]

CONFIG is one of
  1  plain JDQZ, n = 2000 (block size 4, block-tridiagonal, 2 tau rows), k = 8,
     sigma = 0.5+0.5j, maxsub 40, minsub 20, exact LU
  2  plain FP64, n = 1e6 (block size 8, block-pentadiagonal, 16 tau rows),
     k = 64, sigma = 3, maxsub 200, minsub 80, exact bordered banded LU
  3  mixed precision, n = 250000 (block size 8, pentadiagonal, 8 tau rows),
     k = 32, sigma = 1+2j, maxsub 120, minsub 48, inner GMRES(50) at
     inner_tol 1e-3 in complex64, block-Jacobi preconditioner
     (--pc shifted: the exact LU of A - (sigma + delta) B instead, an
     approximate preconditioner as in the reference's own GMRES test, delta
     = --pc-shift, default 0.02); the same solve in FP64 gives the
     eigenvalue error of the mixed run
  5  batched small solves: R (default 2048) pencils, n log-uniform on
     [4096, 32768] (multiple of 4; block size 4, tridiagonal, 2 tau rows),
     k = 8, maxsub 32, minsub 16, sigma uniform on the unit disk, inner
     GMRES(20) at inner_tol 1e-3 in complex64 with block-Jacobi (--pc
     exact: the exact LU at sigma), through the serving loop in batches of
     256 requests

Pencils are the seeded generator's (CN(0,1) entries, +4 on the diagonal
blocks, B = I on the non-tau rows and 0 on the tau rows).  Every converged
pair is checked in FP64: ||A x - lambda B x|| / (||A|| + |lambda| ||B||),
with the device operator's matvecs on the returned eigenvector and ||A||
bounded below by the largest row 2-norm.  One JSON line per run on stdout.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _lib
from .bbjd import jdqz
from .pencil import BandedTauPencil, DeviceOperator


def row_norm_bound(p):
    """largest row 2-norm of A (a lower bound of ||A||_2)"""
    nb, nd, bs = p.n // p.bs, 2 * p.hw + 1, p.bs
    blk = p.band.reshape(nb, nd, bs, bs)
    rows = np.sqrt((np.abs(blk) ** 2).sum(axis=(1, 2))).reshape(-1)[: p.n1]
    m = float(rows.max()) if rows.size else 0.0
    if p.ntau:
        m = max(m, float(np.sqrt((np.abs(p.tau.reshape(p.ntau, p.n)) ** 2).sum(axis=1)).max()))
    return m


def pair_residuals(op, p, evals, X):
    """FP64 residuals of the pairs (columns of X) with the device matvecs"""
    na = row_norm_bound(p)
    nbd = float(np.abs(p.bdiag).max())
    out = []
    for j, lam in enumerate(evals):
        x = np.ascontiguousarray(X[:, j])
        r = op.matvec(x) - lam * op.matvec_B(x)
        out.append(float(np.linalg.norm(r) / ((na + abs(lam) * nbd) * np.linalg.norm(x))))
    return np.asarray(out)


def rel_diff(a, b):
    """max over a of the relative distance to the nearest value of b"""
    if not len(a) or not len(b):
        return None
    return float(max(np.min(np.abs(b - z)) / abs(z) for z in a))


def plain(cfg, n, bs, hw, ntau, k, sigma, maxsub, minsub, pc, seed, inner=0, inner_tol=1e-2,
          precision="fp64", maxiter=2000, verbose=1, pc_shift=0.0, graphs=False, inner_form="reference"):
    t0 = time.perf_counter()
    p = BandedTauPencil.generate(n, bs, hw, ntau, seed=seed)
    op = DeviceOperator(p, sigma + pc_shift, pc=pc, c64=precision == "mixed")
    t_setup = time.perf_counter() - t0
    t = time.perf_counter()
    ev, X, info = jdqz(op, k=k, sigma=sigma, tol=1e-10, maxsub=maxsub, minsub=minsub, maxiter=maxiter,
                       inner_maxiter=inner, inner_tol=inner_tol, preconditioner=op, return_info=True,
                       precision=precision, verbose=verbose, graphs=graphs,
                       inner_form=inner_form if precision == "mixed" else "reference")
    dt = time.perf_counter() - t
    res = pair_residuals(op, p, ev, X) if len(ev) else np.zeros(0)
    return {"config": cfg, "n": n, "k": k, "sigma": [sigma.real, sigma.imag], "maxsub": maxsub,
            "minsub": minsub, "preconditioner": pc + (f" at sigma + {pc_shift}" if pc_shift else ""),
            "inner_maxiter": inner, "precision": precision, "graphs": bool(graphs),
            "inner_form": inner_form if precision == "mixed" else "reference", "_ev": ev,
            "nconv": int(info["nconv"]), "iters": int(info["iters"]), "status": int(info["status"]),
            "solve_seconds": round(dt, 2), "eigenpairs_per_s": round(info["nconv"] / dt, 3),
            "setup_seconds": round(t_setup, 2), "device_bytes": op.device_bytes,
            "max_solver_residual": float(np.max(info["residuals"])) if len(ev) else None,
            "max_fp64_residual": float(res.max()) if len(res) else None,
            "pairs_over_tol": int(np.sum(res > 1e-10)),
            "eigenvalues_head": [[float(z.real), float(z.imag)] for z in ev[:4]]}


def batched(requests, seed, workers=16, pc="block_jacobi", graphs=False):
    from . import serve as srv
    srv.device_init(sync_yield=True)
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(4096), np.log(32768), size=requests))
    ns = (np.round(ns / 4) * 4).astype(int)
    rad = np.sqrt(rng.uniform(0, 1, size=requests))
    ang = rng.uniform(0, 2 * np.pi, size=requests)
    sig = rad * np.exp(1j * ang)
    seeds = rng.integers(1, 2**62, size=requests)
    pencils = [BandedTauPencil.generate(int(ns[i]), 4, 1, 2, seed=int(seeds[i])) for i in range(requests)]
    jobs = [dict(pencil=pencils[i], pc=pc, sigma=complex(sig[i]), nev=8, maxsub=32, minsub=16,
                 maxiter=1000, inner_maxiter=20, inner_tol=1e-3, tol=1e-10) for i in range(requests)]
    # served in batches of 256 requests (a progress line after each)
    res, wall, pairs, used = [], 0.0, 0, 0
    for b0 in range(0, requests, 256):
        t = time.perf_counter()
        r, rep = srv.serve(jobs[b0:b0 + 256], workers=workers, precision="mixed", realtime=False,
                           keep_vectors=True, graphs=graphs)
        dt = time.perf_counter() - t
        res += r
        wall += dt
        pairs += int(rep["pairs"])
        used = max(used, rep["workers"])
        print(f"[configs] batch {b0 // 256}: {rep['pairs']} pairs in {dt:.1f}s", file=sys.stderr, flush=True)
    # FP64 residuals of every pair on a sample of requests (device operators)
    worst, checked, bad = 0.0, 0, 0
    for i in range(0, requests, max(1, requests // 64)):
        r = res[i]
        if r["nconv"] == 0:
            continue
        op = DeviceOperator(pencils[i], jobs[i]["sigma"], pc="none")
        rr = pair_residuals(op, pencils[i], r["eigenvalues"], r["eigenvectors"])
        op.free()
        worst = max(worst, float(rr.max()))
        checked += len(rr)
        bad += int(np.sum(rr > 1e-10))
    statuses = sorted(set(r["status"] for r in res))
    return {"config": 5, "preconditioner": pc, "graphs": bool(graphs), "requests": requests, "workers": used,
            "pairs": pairs,
            "eigenpairs_per_s": round(pairs / wall, 2), "wall_seconds": round(wall, 2),
            "requests_complete": int(sum(r["nconv"] == 8 for r in res)), "statuses": statuses,
            "fp64_residual_pairs_checked": checked, "max_fp64_residual": worst, "pairs_over_tol": bad,
            "latency_p50_s": round(float(np.percentile([r["t_done"] - r["t_admit"] for r in res], 50)), 3),
            "latency_p99_s": round(float(np.percentile([r["t_done"] - r["t_admit"] for r in res], 99)), 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("config", type=int, choices=[1, 2, 3, 5])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--requests", type=int, default=2048)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--verbose", type=int, default=1, help="solver verbosity (stderr)")
    ap.add_argument("--pc", choices=["block_jacobi", "shifted", "exact"], default="block_jacobi",
                    help="preconditioner: config 3 block_jacobi or shifted, config 5 block_jacobi or exact")
    ap.add_argument("--pc-shift", type=float, default=0.02, help="config 3 shifted: delta")
    ap.add_argument("--inner-form", choices=["reference", "projected"], default="reference",
                    help="config 3: the complex64 inner solve's form")
    ap.add_argument("--n", type=int, default=0, help="config 3: a smaller n (0: 250000)")
    ap.add_argument("--graphs", action="store_true",
                    help="configs 3 and 5: replay the inner GMRES steps from captured HIP graphs")
    a = ap.parse_args()
    _lib.load()
    if a.config == 1:
        out = plain(1, 2000, 4, 1, 2, 8, 0.5 + 0.5j, 40, 20, "exact", a.seed, verbose=a.verbose)
        out.pop("_ev")
    elif a.config == 2:
        out = plain(2, 1_000_000, 8, 2, 16, 64, 3.0 + 0j, 200, 80, "exact", a.seed, verbose=a.verbose)
        out.pop("_ev")
    elif a.config == 3:
        shift = a.pc_shift if a.pc == "shifted" else 0.0
        pc = "exact" if a.pc == "shifted" else "block_jacobi"
        n3 = a.n or 250_000
        out = plain(3, n3, 8, 2, 8, 32, 1 + 2j, 120, 48, pc, a.seed, inner=50, inner_tol=1e-3,
                    precision="mixed", verbose=a.verbose, pc_shift=shift, graphs=a.graphs,
                    inner_form=a.inner_form)
        if out["nconv"]:
            ref = plain(3, n3, 8, 2, 8, 32, 1 + 2j, 120, 48, pc, a.seed, inner=50, inner_tol=1e-3,
                        precision="fp64", verbose=a.verbose, pc_shift=shift)
            out["fp64_nconv"] = ref["nconv"]
            out["fp64_solve_seconds"] = ref["solve_seconds"]
            out["max_rel_eig_diff_vs_fp64"] = rel_diff(out["_ev"], ref["_ev"])
            out.pop("_ev")
        else:
            out.pop("_ev")
    else:
        out = batched(a.requests, a.seed, a.workers, pc="exact" if a.pc == "exact" else "block_jacobi",
                      graphs=a.graphs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.exit(main())
