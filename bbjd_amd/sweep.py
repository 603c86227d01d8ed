"""σ sweeps over the serving loop.

The reference runs an eigenspectrum sweep as independent solves, one σ per
MPI rank, and leaves the merging of their results to the user
(README.md:59-97; a start vector "from a coarser discretization or a nearby
parameter", README.md:112-113).  ``sweep`` serves the σ points as requests
on one GPU and adds the two conveniences a sweep needs:

- ``chain=True`` solves the σ points in two waves (even positions of the
  σ-sorted list first).  A second-wave σ that lies within the eigenvalue
  disk of the first-wave request with the nearest σ (its eigenvalues' largest
  distance from that σ) starts from a mix of that request's eigenvectors
  weighted by 1 / |λ − σ|; other σ points start as usual.  Not from one
  exact eigenvector: its residual would be below tol at once and lock an
  eigenvalue that need not be among the nev nearest the new σ;
- the eigenvalues found by several requests (overlapping windows) are merged
  (``dedup``): two eigenvalues within ``rtol · max(1, |λ|)`` of each other
  are one, represented by the pair with the smaller residual.

Each request's own results are returned unchanged next to the merged set.
"""
from __future__ import annotations

import numpy as np

from . import serve as srv


def dedup(eigenvalues, residuals, rtol=1e-8):
    """Indices of the representatives of ``eigenvalues`` after merging those
    within ``rtol · max(1, |λ|)`` of each other (transitively), each group
    represented by its smallest residual; sorted by (real, imaginary) part.
    Returns (keep, groups): keep[g] an index into the input, groups[g] all
    indices merged into it."""
    lam = np.asarray(eigenvalues, dtype=np.complex128).ravel()
    res = np.asarray(residuals, dtype=np.float64).ravel()
    m = lam.size
    if m == 0:
        return np.zeros(0, dtype=int), []
    parent = list(range(m))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    order = np.argsort(lam.real, kind="stable")
    scale = np.maximum(1.0, np.abs(lam))
    for a in range(m):
        i = order[a]
        for b in range(a + 1, m):
            j = order[b]
            tol_ij = rtol * max(scale[i], scale[j])
            if lam[j].real - lam[i].real > tol_ij:
                break
            if abs(lam[j] - lam[i]) <= tol_ij:
                ri, rj = find(i), find(j)
                if ri != rj:
                    parent[rj] = ri
    groups = {}
    for i in range(m):
        groups.setdefault(find(i), []).append(i)
    out = []
    for members in groups.values():
        best = min(members, key=lambda t: (res[t], t))
        out.append((best, sorted(members)))
    out.sort(key=lambda bg: (lam[bg[0]].real, lam[bg[0]].imag))
    return np.array([b for b, _ in out], dtype=int), [g for _, g in out]


def sweep(pencil, sigmas, nev, tol=1e-10, maxsub=0, minsub=0, maxiter=1000, workers=16, pc="exact",
          precision="fp64", chain=False, rtol=1e-8):
    """Solve ``nev`` eigenpairs nearest each σ of ``sigmas`` for one
    ``BandedTauPencil`` through the serving loop and merge the results.

    Returns a dict: ``eigenvalues`` / ``eigenvectors`` (n x m) / ``residuals``
    (‖A x − λ B x‖ with ‖x‖ = 1) of the merged set, ``found_by`` (the
    requests each merged eigenvalue came from), ``requests`` (the serving
    loop's per-request results, in the order of ``sigmas``), ``duplicates``
    (pairs merged away) and ``wall_seconds``."""
    sig = [complex(s) for s in sigmas]
    base = dict(pencil=pencil, pc=pc, nev=int(nev), maxsub=int(maxsub), minsub=int(minsub),
                maxiter=int(maxiter), tol=float(tol))
    srt = sorted(range(len(sig)), key=lambda i: (sig[i].real, sig[i].imag))
    waves = [srt[0::2], srt[1::2]] if chain and len(sig) > 1 else [srt]
    results = [None] * len(sig)
    wall = 0.0
    for w, idx in enumerate(waves):
        jobs = []
        for i in idx:
            job = dict(base, sigma=sig[i])
            if w > 0:
                job["v0"] = _neighbour_start(results, sig, sig[i])
            jobs.append(job)
        res, rep = srv.serve(jobs, workers=min(workers, len(jobs)), precision=precision, realtime=False,
                             keep_vectors=True)
        wall += rep["wall_seconds"]
        for i, r in zip(idx, res):
            results[i] = r
    lam, vec, pres, who = [], [], [], []
    for i, r in enumerate(results):
        for c in range(r["nconv"]):
            lam.append(r["eigenvalues"][c])
            vec.append(r["eigenvectors"][:, c])
            pres.append(r["pair_residuals"][c])
            who.append(i)
    keep, groups = dedup(lam, pres, rtol)
    n = pencil.n
    return {"eigenvalues": np.array([lam[k] for k in keep], dtype=np.complex128),
            "eigenvectors": (np.array([vec[k] for k in keep]).T if len(keep) else np.zeros((n, 0), np.complex128)),
            "residuals": np.array([pres[k] for k in keep]),
            "found_by": [sorted({who[t] for t in g}) for g in groups],
            "requests": results, "duplicates": len(lam) - len(keep), "wall_seconds": wall}


def _neighbour_start(results, sig, s):
    """a start vector for σ = s from the finished request with the nearest
    σ, when s lies within its eigenvalue disk: its eigenvectors weighted by
    1 / |λ - s|, normalised (None otherwise)"""
    done = [i for i, r in enumerate(results) if r is not None and r["nconv"] > 0]
    if not done:
        return None
    i = min(done, key=lambda t: abs(sig[t] - s))
    r = results[i]
    lam = r["eigenvalues"]
    if abs(s - sig[i]) > np.max(np.abs(lam - sig[i])):
        return None
    w = 1.0 / (np.abs(lam - s) + 1e-300)
    v = r["eigenvectors"] @ (w / w.sum())
    nv = np.linalg.norm(v)
    return v / nv if nv > 0 else None
