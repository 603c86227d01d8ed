#!/usr/bin/env python3
"""bench.py -- converged eigenpairs/s at tol 1e-10, served vs plain, one MI355X,
with p50/p99 solve latency.

Workload (one step): the serving-loop stream of 256 JDQZ requests with
Poisson arrivals at a fixed rate; each request draws n from {2e4, 5e4, 1e5,
2e5} (block size 8, block-tridiagonal, 4 dense tau rows, CN(0,1) entries +4
on the diagonal blocks), k from {8, 16, 32}, sigma uniform on [-4,4]+[-4,4]i,
tol 1e-10, maxsub 4k, minsub 2k, exact banded LU of A - sigma B as the
preconditioner.  The operators (pencil + factorisation) are built and resident
in HBM before the timed region, and so are the serving loop's worker threads,
streams and workspaces (one resident server for every step); the timed
region is the serving loop draining the stream.  Timed steps run as arrival
waves of one server run (--pipeline steps per run): a step's requests start
arriving when the previous step's last request has been admitted, so workers
freed in one step's tail take the next step's requests.

  value        served eigenpairs/s over the timed steps (all pairs converged
               at tol).  The served form is the FP64 engine with the
               reference's one-step correction, so each request's results are
               those of the plain path (and of the reference's trajectory);
               --served-precision mixed --served-inner m serves the
               complex64 inner GMRES(m) form instead
  verification every pair returned by the first timed step is checked
               afterwards with a fresh FP64 operator apply on the returned
               vector: ||A x - lambda B x|| / (||A|| + |lambda| ||B||) <= tol,
               ||A|| bounded below by its largest row norm (stricter)
  plain        the same solves one at a time (FP64 engine, no overlap): a
               stratified sample -- per (n, k) class of the workload, three
               of its requests (seeded) solved alone -- gives each class's mean
               pairs and seconds per request; weighted by the class counts of
               the 256 requests they estimate the plain path's eigenpairs/s
               on the whole workload.  The line's estimate weights each
               request by its outer iterations, known exactly from the first
               served step (the served form runs the plain arithmetic, so the
               trajectories are the same; checked on the sample): plain time
               = sum of iterations x the class's sampled seconds per
               iteration.  The class-mean estimate is reported beside it.
               The sampled requests' eigenvalue sets are compared with the
               served ones
  latency      p50/p99 of completion - arrival over the timed requests
  roofline     the projection kernel Q^H x (the dominant streaming kernel):
               durations of every first-pass launch of the plain leg (one
               stream) from start/stop HIP events carried by each dispatch,
               vs its algorithmic bytes 16 n (k + 1), HBM peak 8 TB/s;
               traffic = average algorithmic bytes x the counter-measured
               ratio in profiles/k_proj_traffic.json
  cpu_baseline the reference solver compiled from its sources (oracle/_ref;
               else the oracle restatement) on the host cores, on a bounded
               stratified sample of the workload's requests (one request from
               each of the cheapest (n, k) classes), CPU exact banded LU, all
               cores; also one request on one thread.  Its eigenvalues are
               compared with the served results of the same requests
               (max_rel_eig_err_vs_reference).

--workload small: the batched small solves instead -- 2048 pencils with n
log-uniform on [4096, 32768] (block size 4, block-tridiagonal, 2 tau rows),
k = 8, maxsub 32, minsub 16, sigma uniform on the unit disk, GMRES(20) in
complex64 (the reference's GMRES; --served-inner-form projected for the
projected preconditioner) with the exact banded LU as its preconditioner,
all submitted at once; plain = a random sample of them solved one at a time
in the same form.

Usage: python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--nreq 256]
       python bench.py --plain-only ...   (the plain leg alone: for profilers)
       python bench.py --served-only ...  (the served steps alone: for profilers)
       python bench.py --workload small [--nreq 2048]
One JSON line on stdout; progress on stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
T_START = time.perf_counter()


def log(*a):
    print(f"[bench {time.perf_counter() - T_START:6.1f}s]", *a, file=sys.stderr, flush=True)


def make_requests(nreq, rate, seed):
    rng = np.random.default_rng(seed)
    ns = rng.choice([20000, 50000, 100000, 200000], size=nreq)
    ks = rng.choice([8, 16, 32], size=nreq)
    sig = rng.uniform(-4, 4, size=nreq) + 1j * rng.uniform(-4, 4, size=nreq)
    arr = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    pseeds = rng.integers(1, 2**62, size=nreq)
    return [dict(n=int(ns[i]), k=int(ks[i]), sigma=complex(sig[i]), arrival=float(arr[i]),
                 pseed=int(pseeds[i])) for i in range(nreq)]


def make_small_requests(nreq, seed):
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(4096), np.log(32768), size=nreq))
    ns = (np.round(ns / 4) * 4).astype(int)
    rad = np.sqrt(rng.uniform(0, 1, size=nreq))
    sig = rad * np.exp(1j * rng.uniform(0, 2 * np.pi, size=nreq))
    pseeds = rng.integers(1, 2**62, size=nreq)
    return [dict(n=int(ns[i]), k=8, sigma=complex(sig[i]), arrival=0.0, pseed=int(pseeds[i]))
            for i in range(nreq)]


# per workload: pencil (bs, hw, ntau), solver settings, precision, realtime
# served: the served form's inner solve; plain: the plain leg's (the serving
# workload's plain path is the FP64 engine with the one-step correction)
WORKLOADS = {
    "serving": dict(shape=(8, 1, 4), realtime=True, maxsub=lambda k: 4 * k, minsub=lambda k: 2 * k,
                    workers=16, inner_batch=0,
                    served=dict(inner=0, inner_tol=1e-2, precision="fp64", inner_form="projected"),
                    plain=dict(inner=0, inner_tol=1e-2, precision="fp64", inner_form="reference")),
    "small": dict(shape=(4, 1, 2), realtime=False, maxsub=lambda k: 32, minsub=lambda k: 16,
                  workers=128, inner_batch=16,
                  served=dict(inner=20, inner_tol=1e-3, precision="mixed", inner_form="reference"),
                  plain=dict(inner=20, inner_tol=1e-3, precision="mixed", inner_form="reference")),
}


def class_of(r, workload):
    """(n, k) class of a request; small solves: n is continuous, eight
    log-spaced n classes"""
    if workload == "serving":
        return (r["n"], r["k"])
    return (int(np.log2(r["n"] / 4096) * 8 / 3), r["k"])


def norm_lower_bound(p):
    """max row 2-norm of A: a lower bound of ||A||_2 (stricter residual test)"""
    nb, nd, bs = p.n // p.bs, 2 * p.hw + 1, p.bs
    blk = p.band.reshape(nb, nd, bs, bs)           # [bi, d, c, r]
    rows = np.sqrt((np.abs(blk) ** 2).sum(axis=(1, 2))).reshape(-1)[: p.n1]
    m = rows.max() if rows.size else 0.0
    if p.ntau:
        m = max(m, np.sqrt((np.abs(p.tau.reshape(p.ntau, p.n)) ** 2).sum(axis=1)).max())
    return float(m)


def pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else None


def host_cpu():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


def verify_pairs(results, ops, norms, tol):
    """FP64 residual of every returned pair from a fresh operator apply on
    the returned vector (bbjd_op_matvec_ABr: r = A x - lambda B x and ||r||
    on the device): worst ||r|| / (||A|| + |lambda| ||B||) with ||x|| = 1,
    ||B|| = 1, the count over tol, and the worst ||r|| / ||x|| alone."""
    from bbjd_amd import _lib
    lib = _lib.load()
    worst, worst_abs, bad, checked = 0.0, 0.0, 0, 0
    for i, r in enumerate(results):
        X = r.get("eigenvectors")
        if X is None or r["nconv"] == 0:
            continue
        op = ops[i]
        n = op.n
        dx = _lib.DeviceBuffer(n * 16)
        da, db, dr = _lib.DeviceBuffer(n * 16), _lib.DeviceBuffer(n * 16), _lib.DeviceBuffer(n * 16)
        dn = _lib.DeviceBuffer(8)
        for j in range(r["nconv"]):
            x = np.ascontiguousarray(X[:, j])
            lam = complex(r["eigenvalues"][j])
            xn = float(np.linalg.norm(x))
            if lib.bbjd_memcpy_h2d(dx.ptr, x.ctypes.data, x.nbytes):
                raise RuntimeError("upload failed")
            rc = lib.bbjd_op_matvec_ABr(n, dx.ptr, da.ptr, db.ptr, _lib.zc(1.0), _lib.zc(-lam), dr.ptr, dn.ptr,
                                        op.handle)
            rc |= lib.bbjd_dev_sync()
            if rc:
                raise RuntimeError("verification apply failed")
            rn = float(dn.to_array(np.float64, 1)[0]) / xn
            v = rn / (norms[i] + abs(lam))
            worst = max(worst, v)
            worst_abs = max(worst_abs, rn)
            bad += not (v <= tol)
            checked += 1
        for b in (dx, da, db, dr, dn):
            b.free()
    return worst, bad, checked, worst_abs


def cpu_baseline(reqs, served, wl, workload, classes_wanted, threads, budget_s):
    """Reference solver on the host (oracle/_ref if built, else the oracle);
    FP64 throughout (the reference has no complex64 inner solve).  One
    request from each wanted class (cheapest first) on `threads` threads,
    until budget_s is spent; then one request of the cheapest class on one
    thread.  Returns the baseline dict and per-request eigenvalue errors
    against the served results of the same requests."""
    from bbjd_amd import BandedTauPencil
    from oracle import orc
    which = "ref" if orc.have_ref() else "oracle"
    lib = orc.load_oracle()

    def run(i, thr):
        r = reqs[i]
        lib.orc_set_threads(int(thr))
        p = BandedTauPencil.generate(r["n"], *wl["shape"], seed=r["pseed"])
        ops = orc.CpuPencilOps(p, r["sigma"], 1)
        t = time.perf_counter()
        out = orc.solve(which, p.n, r["k"], r["sigma"], tol=1e-10, maxsub=wl["maxsub"](r["k"]),
                        minsub=wl["minsub"](r["k"]), maxiter=r.get("maxiter", 1000),
                        inner_maxiter=wl["served"]["inner"], inner_tol=wl["served"]["inner_tol"], ops=ops)
        return out, time.perf_counter() - t, p, ops

    def ranks(p, ops, sigma, k, sets):
        """each set's eigenvalues by their rank in the distance to sigma among
        the pencil's k + 24 eigenvalues nearest sigma (ARPACK on the shift-
        and-invert operator (A - sigma B)^-1 B with the CPU banded LU): rank
        list 0 .. k-1 means the k nearest; None when ARPACK fails"""
        from scipy.sparse.linalg import LinearOperator, eigs
        try:
            A, Bm = p.to_scipy()
            op = LinearOperator(A.shape, matvec=lambda x: ops.apply("orc_precond", Bm @ x), dtype=np.complex128)
            mu = eigs(op, k=k + 24, which="LM", return_eigenvectors=False, tol=1e-12, ncv=max(3 * (k + 24), 96))
        except Exception:  # noqa: BLE001
            return [None] * len(sets)
        w = sigma + 1.0 / mu
        w = w[np.argsort(np.abs(w - sigma))]
        return [sorted(int(np.argmin(np.abs(w - e))) for e in np.asarray(ev)) for ev in sets]

    # the first request (in workload order) of each wanted class
    first = {}
    for i, r in enumerate(reqs):
        first.setdefault(class_of(r, workload), i)
    pairs, secs, done, errs = 0, 0.0, [], []
    for key in classes_wanted:
        if secs >= budget_s or key not in first:
            break
        i = first[key]
        out, dt, p, ops = run(i, threads)
        pairs += out["nconv"]
        secs += dt
        done.append((key, i, out["nconv"], dt))
        if served is not None:
            ge = np.asarray(served[i]["eigenvalues"])
            ce = np.asarray(out["evals"])
            if ge.size and ce.size:
                # each served eigenvalue against its nearest reference one
                e = max(float(np.min(np.abs(ce - g)) / abs(g)) for g in ge)
                # which eigenvalues each returned, by distance rank to sigma
                # (outside the timed CPU solve)
                rs, rr = ranks(p, ops, reqs[i]["sigma"], reqs[i]["k"], [ge, ce]) if p.n <= 60000 else (None, None)
                errs.append((i, e, len(ge), len(ce), rs, rr))
    one = None
    if done:
        i = done[0][1]          # the cheapest class's request
        out1, dt1, _, _ = run(i, 1)
        one = {"value": out1["nconv"] / dt1, "cores": 1, "sample": f"request {i} (n={reqs[i]['n']}, "
               f"k={reqs[i]['k']}): {out1['nconv']} pairs in {dt1:.2f} s"}
    model, avail = host_cpu()
    base = {"value": pairs / secs if secs else None, "unit": "eigenpairs/s", "cores": int(threads),
            "kind": "reference" if which == "ref" else "port",
            "sample": (f"one request from each of the {len(done)} cheapest (n, k) classes of the workload "
                       f"(first in workload order), solved to tol 1e-10 with a CPU exact banded LU (LAPACK "
                       f"zgbtrf + Schur complement of the tau rows)"
                       + (f" as the preconditioner of an FP64 inner GMRES({wl['served']['inner']}) (the served "
                          f"form's inner solve; the reference has no complex64 one)" if wl["served"]["inner"] else "")
                       + ": "
                       + "; ".join(f"n={k[0]} k={k[1]}: {c} pairs {d:.1f}s" for k, _, c, d in done)),
            "classes": [list(k) for k, _, _, _ in done],
            "one_thread": one, "host_cpu": model, "host_cpus_available": avail}
    return base, errs, done


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["serving", "small"], default="serving")
    ap.add_argument("--nreq", type=int, default=0, help="requests (default 256 serving, 2048 small)")
    ap.add_argument("--rate", type=float, default=32.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--workers", type=int, default=0,
                    help="serving workers (default: 16 serving, 128 small, whose inner solves run grouped)")
    ap.add_argument("--maxiter", type=int, default=1000)
    ap.add_argument("--plain-per-class", type=int, default=3, help="plain solves per (n, k) class")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-verify", action="store_true", help="skip the residual check of the returned pairs")
    ap.add_argument("--plain-only", action="store_true", help="skip the served steps (profiling the plain leg)")
    ap.add_argument("--served-only", action="store_true", help="skip the plain leg (profiling the served steps)")
    ap.add_argument("--no-tile-table", action="store_true", help="engine default tiles instead of the autotuned table")
    ap.add_argument("--order", choices=["arrival", "largest", "window"], default="arrival",
                    help="which arrived request a free worker takes (serving loop)")
    ap.add_argument("--window", type=int, default=8,
                    help="--order window: the largest of this many longest-waiting requests")
    ap.add_argument("--sync", choices=["yield", "block"], default="block",
                    help="how the served leg's workers wait on the GPU (block: each workspace's waits sleep "
                         "until the GPU signals; the plain leg always yields)")
    ap.add_argument("--pin", action="store_true", help="pin each serving worker to one CPU")
    ap.add_argument("--gmres-lookahead", type=int, default=4,
                    help="Arnoldi steps of the inner GMRES in flight (1: one at a time)")
    ap.add_argument("--proj-blocks", type=int, default=0,
                    help="blocks per projection launch for every shape (0: the tile table or the engine default)")
    ap.add_argument("--mid-blocks", type=int, default=0,
                    help="largest grid of the CGS2 middle pass (0: engine default)")
    ap.add_argument("--fuse-blocks", type=int, default=0,
                    help="finish reductions in the last block for grids up to this many blocks (0: engine default)")
    ap.add_argument("--operators", choices=["resident", "on-demand"], default="resident",
                    help="on-demand: the server builds each request's operator (host factorization, upload) "
                         "when it admits the request, inside the timed region, charged against free HBM")
    ap.add_argument("--qz-mode", choices=["ritz", "schur", "device"], default="ritz",
                    help="where the projected step runs (device: on the GPU for subspaces up to 64)")
    ap.add_argument("--time-budget", type=float, default=540.0,
                    help="seconds for the whole run: timed steps that would end past it, less what the "
                         "plain leg, the verification and the CPU baseline need, are not started (the "
                         "line then reports the steps timed)")
    ap.add_argument("--served-inner", type=int, default=-1,
                    help="inner GMRES steps of the served form (-1: the workload's default)")
    ap.add_argument("--served-inner-tol", type=float, default=0.0,
                    help="inner GMRES relative tolerance of the served form (0: the workload's default)")
    ap.add_argument("--served-precision", choices=["fp64", "mixed"], default=None,
                    help="inner-solve precision of the served form (default: the workload's)")
    ap.add_argument("--served-inner-form", choices=["reference", "projected"], default=None,
                    help="form of the served form's mixed inner solve (default: the workload's)")
    ap.add_argument("--pipeline", type=int, default=10,
                    help="timed steps per server run, each one arrival wave (1: every step drains before "
                         "the next starts arriving)")
    ap.add_argument("--inner-batch", type=int, default=-1,
                    help="served leg, mixed inner solves: up to this many per grouped batch (0: off; default: "
                         "16 for the small workload, off for the serving workload)")
    ap.add_argument("--inner-batchers", type=int, default=4, help="batcher threads (streams) with --inner-batch")
    ap.add_argument("--inner-window-us", type=float, default=300.0,
                    help="how long a batcher waits for more inner solves (microseconds)")
    ap.add_argument("--batch-pick", choices=["nearest", "arrival"], default="nearest",
                    help="which waiting inner solves a batch takes with the oldest: nearest in n, or arrival order")
    ap.add_argument("--seed", type=int, default=20261203)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solves of this workload are independent and "
              "would be placed per GPU without communication; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from bbjd_amd import BandedTauPencil, DeviceOperator, _lib
    from bbjd_amd import serve as srv

    srv.device_init(sync_yield=True)
    # the committed per-shape tile table of this device, when there is one
    from bbjd_amd import autotune
    tt = None if args.no_tile_table else autotune.apply()
    _lib.load().bbjd_set_gmres_lookahead(args.gmres_lookahead)
    if args.fuse_blocks > 0:
        _lib.load().bbjd_set_fuse_blocks(args.fuse_blocks)
    _lib.load().bbjd_set_batch_pick(1 if args.batch_pick == "nearest" else 0)
    if args.mid_blocks > 0:
        _lib.load().bbjd_set_cgs_mid_blocks(args.mid_blocks)
    if args.proj_blocks > 0:
        _lib.load().bbjd_set_tile_table(0, None, 0, None, None, None)
        _lib.load().bbjd_set_tiles(args.proj_blocks, 0, 0)
    log(f"tile table: {'autotuned ' + tt['tuned'] if tt else 'engine defaults'}")
    wl = dict(WORKLOADS[args.workload])
    wl["served"] = dict(wl["served"])
    if args.served_inner >= 0:
        wl["served"]["inner"] = args.served_inner
    if args.served_inner_tol > 0:
        wl["served"]["inner_tol"] = args.served_inner_tol
    if args.served_precision:
        wl["served"]["precision"] = args.served_precision
    if args.served_inner_form:
        wl["served"]["inner_form"] = args.served_inner_form
    if args.workload == "small":
        # the small solves' plain leg is the served form solved one at a time
        wl["plain"] = dict(wl["served"])
    if args.workers <= 0:
        args.workers = wl["workers"]
    if args.inner_batch < 0:
        args.inner_batch = wl["inner_batch"]
    if wl["served"]["precision"] != "mixed" or not wl["served"]["inner"]:
        args.inner_batch = 0            # only mixed inner solves are grouped
    if args.nreq <= 0:
        args.nreq = 256 if args.workload == "serving" else 2048
    reqs = (make_requests(args.nreq, args.rate, args.seed) if args.workload == "serving"
            else make_small_requests(args.nreq, args.seed))
    for r in reqs:
        r["maxiter"] = args.maxiter

    # ---- resident operators (not timed) ----
    t0 = time.perf_counter()
    ops, norms, pencils = [], [], []
    for i, r in enumerate(reqs):
        p = BandedTauPencil.generate(r["n"], *wl["shape"], seed=r["pseed"])
        norms.append(norm_lower_bound(p))
        op = DeviceOperator(p, r["sigma"], pc="exact", c64="mixed" in (wl["served"]["precision"],
                                                                      wl["plain"]["precision"]))
        op.pencil = None          # host copy no longer needed
        ops.append(op)
        # on-demand operators: the server factors and uploads each request's
        # operator when it admits the request, inside the timed region
        pencils.append(p if args.operators == "on-demand" else None)
    t_setup = time.perf_counter() - t0
    op_bytes = sum(o.device_bytes for o in ops)
    log(f"{len(ops)} operators resident ({op_bytes / 1e9:.1f} GB) in {t_setup:.1f}s; "
        f"free HBM {srv.device_free_bytes() / 1e9:.1f} GB")

    def jobs_for(idx, with_arrival=True, on_demand=False, form="served", wave=0, keep=False):
        src = (lambda i: dict(pencil=pencils[i], pc="exact")) if on_demand else (lambda i: dict(op=ops[i]))
        f = wl[form]
        return [dict(**src(i), sigma=reqs[i]["sigma"], nev=reqs[i]["k"], maxsub=wl["maxsub"](reqs[i]["k"]),
                     minsub=wl["minsub"](reqs[i]["k"]), maxiter=reqs[i]["maxiter"], tol=1e-10,
                     inner_maxiter=f["inner"], inner_tol=f["inner_tol"],
                     arrival=reqs[i]["arrival"] if with_arrival else 0.0, wave=wave, keep_vectors=keep)
                for i in idx]

    all_idx = list(range(len(reqs)))
    served, walls, run_steps = [], [], []
    batch = None
    steps_timed = 0
    worst = bad = checked = worst_abs = None
    if not args.plain_only:
        # ---- the resident server: workers, streams, workspaces (not timed) ----
        t = time.perf_counter()
        server = srv.Server(jobs_for(all_idx), workers=args.workers, precision=wl["served"]["precision"],
                            inner_form=wl["served"]["inner_form"],
                            sync_yield="block" if args.sync == "block" else True, pin=args.pin,
                            qz_mode=args.qz_mode, inner_batch=args.inner_batch,
                            inner_batchers=args.inner_batchers, inner_window_us=args.inner_window_us)
        log(f"server: {server.workers} workers, {server.workspace_bytes / 1e9:.2f} GB workspace each, "
            f"set up in {time.perf_counter() - t:.1f}s; free HBM {srv.device_free_bytes() / 1e9:.1f} GB")
        # ---- warmup (untimed): the smallest requests ----
        small_idx = sorted(all_idx, key=lambda i: (reqs[i]["k"], reqs[i]["n"]))[: args.workers]
        for w in range(args.warmup):
            server.run(jobs_for(small_idx, False), realtime=False)
            log(f"warmup step {w} done")
        # ---- timed: served ----
        # what the legs after the timed steps take (measured: plain 26 s with three solves per class, the
        # CPU baseline up to its budget plus one request of ~23 s, the
        # verification 2 s), with a margin
        reserve = (0 if args.served_only else 35.0) + (0 if args.no_cpu_baseline else args.cpu_seconds + 25.0) \
            + (0 if args.no_verify else 6.0)
        # Steps run in runs of up to --pipeline steps, each step one arrival
        # wave of the run: a step's requests start arriving when the last
        # request of the step before has been admitted, so the workers that
        # free up in one step's tail take the next step's requests (the
        # continuous stream of a server under load) instead of idling until
        # the whole step has finished.  Every request of every timed step is
        # solved inside the timed region.
        nsteps_req = len(all_idx)
        s = 0
        per_step = None
        while s < args.steps:
            elapsed = time.perf_counter() - T_START
            c = min(args.pipeline, args.steps - s)
            if per_step is not None:
                fit = int((args.time_budget - elapsed - reserve) // per_step)
                if fit < c:
                    log(f"time budget: {c - max(fit, 0)} of the remaining {args.steps - s} steps not started "
                        f"({elapsed:.0f}s elapsed, {per_step:.0f}s per step, {reserve:.0f}s reserved)")
                    c = fit
                if c <= 0:
                    break
            jobs = []
            for w in range(c):
                jobs += jobs_for(all_idx, on_demand=args.operators == "on-demand", wave=w,
                                 keep=(s + w == 0 and not args.no_verify))
            free0 = srv.device_free_bytes()
            t = time.perf_counter()
            res, rep = server.run(jobs, realtime=wl["realtime"], order=args.order, window=args.window,
                                  progress=True)
            walls.append(time.perf_counter() - t)
            run_steps.append(c)
            per_step = walls[-1] / c
            for w in range(c):
                rw = res[w * nsteps_req:(w + 1) * nsteps_req]
                served.append((rw, {"pairs": sum(r["nconv"] for r in rw)}))
            steps_timed += c
            log(f"steps {s}-{s + c - 1}: {rep['pairs']} pairs in {walls[-1]:.1f}s ({per_step:.1f}s per step), "
                f"workers {rep['workers']}, peak concurrency {rep['peak_concurrency']}, peak footprint "
                f"{rep['peak_bytes'] / 1e9:.1f} GB, failed {rep['failed']}; free HBM "
                f"{free0 / 1e9:.1f} -> {srv.device_free_bytes() / 1e9:.1f} GB")
            s += c
        batch = server.batch_stats() if args.inner_batch else None
        if batch:
            log(f"grouped inner solves: {batch['jobs']} in {batch['batches']} batches, {batch['steps']} steps, "
                f"{batch['recorded']} launches recorded, {batch['launches']} made")
        server.close()
        if served and not args.no_verify:
            t = time.perf_counter()
            worst, bad, checked, worst_abs = verify_pairs(served[0][0], ops, norms, 1e-10)
            for r in served[0][0]:
                r.pop("eigenvectors", None)
            log(f"verified {checked} pairs of step 0 by fresh operator applies in {time.perf_counter() - t:.1f}s: "
                f"worst ||Ax - lBx|| / (||A|| + |l| ||B||) {worst:.2e}, worst ||Ax - lBx|| {worst_abs:.2e}, "
                f"over tol {bad}")

    pairs = sum(rep["pairs"] for _, rep in served)
    wall = sum(walls)
    # the engine's algorithmic HBM bytes (its n-sized kernels; the operators'
    # applies are not counted) over the served region, and the solves in
    # flight on average
    served_region = None
    if served and wall:
        sb = sum(r["stream_bytes"] for res, _ in served for r in res)
        busy_all = sum(r["t_done"] - r["t_start"] for res, _ in served for r in res)
        served_region = {"engine_stream_GBs": round(sb / wall / 1e9, 1),
                         "frac_of_hbm_peak": round(sb / wall / 1e9 / HBM_PEAK_GBS, 4),
                         "mean_solves_in_flight": round(busy_all / wall, 2)}
    lat = [r["t_done"] - r["t_arrive"] for res, _ in served for r in res]
    statuses = sorted(set(r["status"] for res, _ in served for r in res))
    conv_over_tol = sum(int(np.sum(r["residuals"] >= 1e-10)) for res, _ in served for r in res)
    if served:
        log(f"served: {pairs} pairs / {wall:.1f}s; statuses {statuses}; solver residuals over tol {conv_over_tol}")

    # ---- plain: one solve at a time, FP64, stratified by (n, k) class ----
    rng = np.random.default_rng(args.seed + 1)
    classes = {}
    for i, r in enumerate(reqs):
        classes.setdefault(class_of(r, args.workload), []).append(i)
    plain_rate, per_cls, diff, sample, kcnt, nconv_mismatch = None, {}, None, [], 0, None
    plain_class_mean, plain_estimator = None, None
    if not args.served_only:
        cls_of = {}
        for key in sorted(classes):
            members = classes[key]
            pick = rng.choice(len(members), size=min(args.plain_per_class, len(members)), replace=False)
            for j in sorted(pick.tolist()):
                sample.append(members[j])
                cls_of[members[j]] = key
        srv.ktime_enable(True)
        t = time.perf_counter()
        pres, prep = srv.serve(jobs_for(sample, False, form="plain"), workers=1, realtime=False,
                               precision=wl["plain"]["precision"], inner_form=wl["plain"]["inner_form"],
                               qz_mode=args.qz_mode)
        t_plain = time.perf_counter() - t
        srv.ktime_enable(False)
        kcnt, kms, kbytes = srv.ktime_harvest()
        for j, i in enumerate(sample):
            r = pres[j]
            dt = r["t_done"] - r["t_start"]
            per_cls.setdefault(cls_of[i], []).append((r["nconv"], dt, r["iters"]))
            log(f"  plain n={reqs[i]['n']} k={reqs[i]['k']}: {r['nconv']} pairs, {r['iters']} iters, "
                f"{dt:.2f}s (qz {r['qz_seconds']:.2f}s), status {r['status']}")
        # class means of the sampled solves, weighted by the class counts
        est_pairs = sum(len(classes[c]) * np.mean([v[0] for v in vs]) for c, vs in per_cls.items())
        est_secs = sum(len(classes[c]) * np.mean([v[1] for v in vs]) for c, vs in per_cls.items())
        plain_class_mean = est_pairs / est_secs if est_secs else None
        plain_rate, plain_estimator = plain_class_mean, "class means of the sampled solves, weighted by class counts"
        if served:
            # Iteration-weighted: a request's plain time is its outer
            # iterations times its class's seconds per iteration (from the
            # sampled solves).  The served form runs the plain form's
            # arithmetic, so every request's iterations and pairs are known
            # exactly from the first served step (checked on the sample).
            first = served[0][0]
            iters_match = all(pres[j]["iters"] == first[i]["iters"] and pres[j]["nconv"] == first[i]["nconv"]
                              for j, i in enumerate(sample))
            if iters_match:
                spi = {c: sum(v[1] for v in vs) / max(1, sum(v[2] for v in vs)) for c, vs in per_cls.items()}
                pred_secs = sum(first[i]["iters"] * spi[class_of(reqs[i], args.workload)]
                                for i in range(len(first)) if class_of(reqs[i], args.workload) in spi)
                pred_pairs = sum(first[i]["nconv"] for i in range(len(first))
                                 if class_of(reqs[i], args.workload) in spi)
                if pred_secs > 0:
                    plain_rate = pred_pairs / pred_secs
                    plain_estimator = ("every request's outer iterations (from the first served step, the plain "
                                       "form's trajectory) times its class's plain seconds per iteration from "
                                       "the sampled solves")
            log(f"plain estimators: iteration-weighted {plain_rate:.2f}, class means {plain_class_mean:.2f} "
                f"pairs/s; sampled iterations equal the served ones: {iters_match}")
        if served:
            # the same eigenvalue sets from the plain leg (alone) and the
            # served form (in the mix): each value against its nearest in the
            # other set, both ways; a count mismatch is reported as such
            mix = served[0][0]
            diff, nconv_mismatch = 0.0, 0
            for j, i in enumerate(sample):
                a, b = np.asarray(pres[j]["eigenvalues"]), np.asarray(mix[i]["eigenvalues"])
                if a.size != b.size:
                    nconv_mismatch += 1
                if a.size and b.size:
                    d_ab = max(float(np.min(np.abs(b - x)) / abs(x)) for x in a)
                    d_ba = max(float(np.min(np.abs(a - x)) / abs(x)) for x in b)
                    diff = max(diff, d_ab, d_ba)
            s_alone = sum(pres[j]["t_done"] - pres[j]["t_start"] for j in range(len(sample)))
            s_mix = sum(mix[i]["t_done"] - mix[i]["t_start"] for i in sample)
            q_alone = sum(pres[j]["qz_seconds"] for j in range(len(sample)))
            q_mix = sum(mix[i]["qz_seconds"] for i in sample)
            first_run = [r for res_s, _ in served[:run_steps[0]] for r in res_s]
            busy = sum(r["t_done"] - r["t_start"] for r in first_run)
            log(f"sample solve seconds alone {s_alone:.1f} / in the mix {s_mix:.1f} "
                f"(qz {q_alone:.1f} / {q_mix:.1f}); mean solves in flight {busy / walls[0]:.1f}; "
                f"engine stream {sum(r['stream_bytes'] for r in first_run) / walls[0] / 1e9:.0f} GB/s over the "
                f"first run of {run_steps[0]} steps")
        log(f"plain: {len(sample)} solves in {len(per_cls)} classes, {prep['pairs']} pairs / {t_plain:.1f}s; "
            f"estimate for the workload {plain_rate:.2f} pairs/s; "
            f"max rel eig diff plain vs served {diff}, pair-count mismatches {nconv_mismatch}")

    served_rate = pairs / wall if wall else None
    roof = None
    if kcnt:
        achieved = kbytes / (kms * 1e-3) / 1e9
        # HBM bytes per launch: the counter-measured traffic/algorithmic ratio
        # of this kernel (rocprofv3 FETCH_SIZE + WRITE_SIZE at the workload's
        # average shape, calibrated on a copy kernel) times this run's
        # average algorithmic bytes
        traffic = None
        tj = os.path.join(ROOT, "profiles", "k_proj_traffic.json")
        if os.path.exists(tj):
            with open(tj) as f:
                traffic = round(kbytes / kcnt * json.load(f)["traffic_over_algorithmic"])
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_proj<double,1,16,false> (Q^H x, complex128)", "measured_on": "plain leg",
                "launches": kcnt, "avg_us": round(kms * 1e3 / kcnt, 2), "avg_bytes": round(kbytes / kcnt)}

    cpu, eig_err, sets = None, None, None
    if not args.no_cpu_baseline and not args.plain_only and not args.served_only:
        thr = args.cpu_threads or min(16, host_cpu()[1])
        # the cheapest classes for the CPU first: its cost grows with n and,
        # through the projected QZ, with k
        order = sorted(per_cls, key=lambda c: (c[0] * c[1], c))
        try:
            cpu, errs, done = cpu_baseline(reqs, served[0][0] if served else None, wl, args.workload, order,
                                           thr, args.cpu_seconds)
            if errs:
                eig_err = max(e[1] for e in errs)
                for i, e, ng, nc, rs, rr in errs:
                    log(f"  request {i}: served {ng} / reference {nc} eigenvalues, max rel diff {e:.2e}; "
                        f"distance ranks served {rs}, reference {rr}")
                knear = lambda rk, k: rk is not None and rk == list(range(k))
                sets = {"requests": len(errs),
                        "equal_within_1e-8": sum(e[1] <= 1e-8 for e in errs),
                        "served_are_the_k_nearest": sum(knear(e[4], reqs[e[0]]["k"]) for e in errs),
                        "reference_are_the_k_nearest": sum(knear(e[5], reqs[e[0]]["k"]) for e in errs),
                        "ranks": [{"request": e[0], "served": e[4], "reference": e[5]} for e in errs]}
            # the plain GPU leg's rate on the same classes, for a like-for-like ratio
            if done:
                gp = sum(np.mean([v[0] for v in per_cls[k]]) for k, _, _, _ in done)
                gs = sum(np.mean([v[1] for v in per_cls[k]]) for k, _, _, _ in done)
                cpu["gpu_plain_on_sample_classes"] = round(gp / gs, 3) if gs else None
            log(f"cpu baseline: {cpu['value']} eigenpairs/s ({cpu['kind']}, {cpu['cores']} threads, "
                f"{cpu['host_cpu']}); one thread {cpu['one_thread']}; max rel eig err served vs reference "
                f"{eig_err} over {len(errs)} requests")
        except Exception as e:  # noqa: BLE001
            log(f"cpu baseline unavailable: {e}")

    out = {
        "metric": "eigenpairs/s @tol1e-10 (served vs plain, 1 GPU); p50/p99 solve latency",
        "value": round(served_rate, 3) if served_rate else None,
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": steps_timed,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / steps_timed, 1) if steps_timed else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c128 (outer loop), c64 (inner solve)" if wl["served"]["precision"] == "mixed"
                  and wl["served"]["inner"] else "c128"),
        "data": "synthetic",
        "config": {"workload": (f"serving loop: {args.nreq} JDQZ requests, Poisson arrivals "
                                f"{args.rate}/s, n in {{2e4,5e4,1e5,2e5}} (bs 8 block-tridiagonal, "
                                f"4 tau rows), k in {{8,16,32}}, sigma in [-4,4]+[-4,4]i, tol 1e-10, "
                                f"maxsub 4k, minsub 2k, maxiter {args.maxiter}, exact banded LU"
                                if args.workload == "serving" else
                                f"batched small solves: {args.nreq} JDQZ requests submitted at once, n "
                                f"log-uniform on [4096, 32768] (bs 4 block-tridiagonal, 2 tau rows), k 8, "
                                f"sigma on the unit disk, tol 1e-10, maxsub 32, minsub 16, maxiter "
                                f"{args.maxiter}, GMRES(20) inner solve in complex64 (inner tol 1e-3) with "
                                f"the exact banded LU as its preconditioner"),
                   "workers": args.workers, "requests": args.nreq,
                   "admission": args.order + (f" {args.window}" if args.order == "window" else ""),
                   "host_wait": args.sync, "workers_pinned": bool(args.pin),
                   "projected_step": args.qz_mode, "operators": args.operators,
                   "gmres_lookahead": args.gmres_lookahead, "steps_per_run": args.pipeline},
        "steps_requested": args.steps,
        "served_eigenpairs_per_s": round(served_rate, 3) if served_rate else None,
        "plain_eigenpairs_per_s": round(plain_rate, 3) if plain_rate else None,
        "served_over_plain": round(served_rate / plain_rate, 2) if served_rate and plain_rate else None,
        "plain_sample_requests": len(sample),
        "plain_estimate": plain_estimator if not args.served_only else None,
        "plain_class_mean_eigenpairs_per_s": round(plain_class_mean, 3) if plain_class_mean else None,
        "latency_p50_s": round(pct(lat, 50), 3) if lat else None,
        "latency_p99_s": round(pct(lat, 99), 3) if lat else None,
        "pairs": pairs,
        "statuses": statuses,
        "solver_residuals_over_tol": conv_over_tol,
        "verified_pairs": checked,
        "verified_max_normalised_residual": worst,
        "verified_max_residual": worst_abs,
        "verified_pairs_over_tol": bad,
        "max_rel_eig_err_vs_reference": eig_err,
        "eigenvalue_sets_vs_reference": sets,
        "max_rel_eig_diff_plain_vs_served": diff,
        "plain_vs_served_pair_count_mismatches": nconv_mismatch,
        "served_form": (f"{wl['served']['precision']} inner GMRES({wl['served']['inner']}), inner tol "
                        f"{wl['served']['inner_tol']:g}, {wl['served']['inner_form']} form" if wl["served"]["inner"] else
                        f"{wl['served']['precision']}, one-step correction"),
        "plain_form": (f"{wl['plain']['precision']} inner GMRES({wl['plain']['inner']}), inner tol "
                       f"{wl['plain']['inner_tol']:g}" if wl["plain"]["inner"] else
                       f"{wl['plain']['precision']}, one-step correction"),
        "inner_batching": (dict(max_jobs=args.inner_batch, batchers=args.inner_batchers,
                                window_us=args.inner_window_us, pick=args.batch_pick, **batch)
                           if args.inner_batch and batch else None),
        "setup_seconds": round(t_setup, 1),
        "roofline": roof,
        "served_region": served_region,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
