#!/usr/bin/env python3
"""bench.py -- converged eigenpairs/s at tol 1e-10, served vs plain, one MI355X,
with p50/p99 solve latency.

Workload (one step): the serving-loop stream of 256 JDQZ requests with
Poisson arrivals at a fixed rate; each request draws n from {2e4, 5e4, 1e5,
2e5} (block size 8, block-tridiagonal, 4 dense tau rows, CN(0,1) entries +4
on the diagonal blocks), k from {8, 16, 32}, sigma uniform on [-4,4]+[-4,4]i,
tol 1e-10, maxsub 4k, minsub 2k, exact banded LU of A - sigma B as the
preconditioner.  The operators (pencil + factorisation) are built and resident
in HBM before the timed region; the timed region is the serving loop.

  value        served eigenpairs/s over the timed steps (all pairs converged
               at tol, each pair's FP64 residual checked)
  plain        the same solves one at a time (FP64 engine, no overlap): a
               stratified sample -- per (n, k) class of the workload, two of
               its requests (seeded) solved alone -- gives each class's mean
               pairs and seconds per request; weighted by the class counts of
               the 256 requests they estimate the plain path's eigenpairs/s
               on the whole workload
  latency      p50/p99 of completion - arrival over the timed requests
  roofline     the projection kernel Q^H x (the dominant streaming kernel):
               durations of every first-pass launch of the plain leg (one
               stream) from start/stop HIP events carried by each dispatch,
               vs its algorithmic bytes 16 n (k + 1), HBM peak 8 TB/s; the
               same kernel's rate inside the served region (launches
               time-share the GPU there) is reported beside it; traffic =
               average algorithmic bytes x the counter-measured ratio in
               profiles/k_proj_traffic.json
  cpu_baseline the reference solver compiled from its sources (oracle/_ref;
               else the oracle restatement) on the host cores, on a bounded
               sample of the workload's requests, CPU exact banded LU

--workload small: the batched small solves instead -- 2048 pencils with n
log-uniform on [4096, 32768] (block size 4, block-tridiagonal, 2 tau rows),
k = 8, maxsub 32, minsub 16, sigma uniform on the unit disk, GMRES(20) in
complex64 with the exact banded LU as its preconditioner, all submitted at
once; plain = a random sample of them solved one at a time.

Usage: python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--nreq 256]
       python bench.py --plain-only ...   (the plain leg alone: for profilers)
       python bench.py --workload small [--nreq 2048]
One JSON line on stdout; progress on stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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
WORKLOADS = {
    "serving": dict(shape=(8, 1, 4), inner=0, inner_tol=1e-2, precision="fp64", realtime=True,
                    maxsub=lambda k: 4 * k, minsub=lambda k: 2 * k),
    "small": dict(shape=(4, 1, 2), inner=20, inner_tol=1e-3, precision="mixed", realtime=False,
                  maxsub=lambda k: 32, minsub=lambda k: 16),
}


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


def cpu_baseline(reqs, budget_s, threads, wl):
    """Reference solver on the host (oracle/_ref if built, else the oracle);
    FP64 throughout (the reference has no complex64 inner solve)."""
    from bbjd_amd import BandedTauPencil
    from oracle import orc
    which = "ref" if orc.have_ref() else "oracle"
    lib = orc.load_oracle()
    lib.orc_set_threads(int(threads))
    pairs, secs, done = 0, 0.0, []
    # smallest requests first, so the bounded sample finishes a few solves
    for r in sorted(reqs, key=lambda r: (r["n"], r["k"])):
        if secs >= budget_s:
            break
        p = BandedTauPencil.generate(r["n"], *wl["shape"], seed=r["pseed"])
        ops = orc.CpuPencilOps(p, r["sigma"], 1)
        t = time.perf_counter()
        out = orc.solve(which, p.n, r["k"], r["sigma"], tol=1e-10, maxsub=wl["maxsub"](r["k"]),
                        minsub=wl["minsub"](r["k"]), maxiter=r.get("maxiter", 1000), inner_maxiter=wl["inner"],
                        inner_tol=wl["inner_tol"], ops=ops)
        dt = time.perf_counter() - t
        pairs += out["nconv"]
        secs += dt
        done.append(f"n={r['n']} k={r['k']}: {out['nconv']} pairs {dt:.1f}s")
    return {"value": pairs / secs if secs else None, "unit": "eigenpairs/s", "cores": int(threads),
            "kind": "reference" if which == "ref" else "port",
            "sample": f"{len(done)} of the workload's requests, smallest first, solved to tol 1e-10 "
                      f"with a CPU exact banded LU (LAPACK zgbtrf + Schur complement of the tau rows): "
                      + "; ".join(done)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["serving", "small"], default="serving")
    ap.add_argument("--nreq", type=int, default=0, help="requests (default 256 serving, 2048 small)")
    ap.add_argument("--rate", type=float, default=32.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--maxiter", type=int, default=1000)
    ap.add_argument("--plain-per-class", type=int, default=2, help="plain solves per (n, k) class")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-only", action="store_true", help="skip the served steps (profiling the plain leg)")
    ap.add_argument("--no-tile-table", action="store_true", help="engine default tiles instead of the autotuned table")
    ap.add_argument("--order", choices=["arrival", "largest", "window"], default="arrival",
                    help="which arrived request a free worker takes (serving loop)")
    ap.add_argument("--window", type=int, default=8,
                    help="--order window: the largest of this many longest-waiting requests")
    ap.add_argument("--sync", choices=["yield", "block"], default="yield",
                    help="how host threads wait on the GPU")
    ap.add_argument("--pin", action="store_true", help="pin each serving worker to one CPU")
    ap.add_argument("--seed", type=int, default=20261203)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solves of this workload are independent and "
              "would be placed per GPU without communication; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from bbjd_amd import BandedTauPencil, DeviceOperator
    from bbjd_amd import serve as srv

    srv.device_init(sync_yield="block" if args.sync == "block" else True)
    # the committed per-shape tile table of this device, when there is one
    from bbjd_amd import autotune
    tt = None if args.no_tile_table else autotune.apply()
    log(f"[bench] tile table: {'autotuned ' + tt['tuned'] if tt else 'engine defaults'}")
    wl = WORKLOADS[args.workload]
    if args.nreq <= 0:
        args.nreq = 256 if args.workload == "serving" else 2048
    reqs = (make_requests(args.nreq, args.rate, args.seed) if args.workload == "serving"
            else make_small_requests(args.nreq, args.seed))
    for r in reqs:
        r["maxiter"] = args.maxiter

    # ---- resident operators (not timed) ----
    t0 = time.perf_counter()
    ops, norms = [], []
    for i, r in enumerate(reqs):
        p = BandedTauPencil.generate(r["n"], *wl["shape"], seed=r["pseed"])
        norms.append(norm_lower_bound(p))
        op = DeviceOperator(p, r["sigma"], pc="exact", c64=wl["precision"] == "mixed")
        op.pencil = None          # host copy no longer needed
        ops.append(op)
    t_setup = time.perf_counter() - t0
    op_bytes = sum(o.device_bytes for o in ops)
    log(f"[bench] {len(ops)} operators resident ({op_bytes / 1e9:.1f} GB) in {t_setup:.1f}s")

    def jobs_for(idx, with_arrival=True):
        return [dict(op=ops[i], sigma=reqs[i]["sigma"], nev=reqs[i]["k"], maxsub=wl["maxsub"](reqs[i]["k"]),
                     minsub=wl["minsub"](reqs[i]["k"]), maxiter=reqs[i]["maxiter"], tol=1e-10,
                     inner_maxiter=wl["inner"], inner_tol=wl["inner_tol"],
                     arrival=reqs[i]["arrival"] if with_arrival else 0.0) for i in idx]

    # ---- warmup (untimed): short streams ----
    for w in range(args.warmup):
        idx = list(range(min(16, len(reqs))))
        srv.serve(jobs_for(idx, False), workers=1 if args.plain_only else args.workers, realtime=False,
                  precision=wl["precision"])
        log(f"[bench] warmup step {w} done")

    # ---- timed: served ----
    all_idx = list(range(len(reqs)))
    served, walls = [], []
    srv.ktime_enable(True)
    for s in range(0 if args.plain_only else args.steps):
        t = time.perf_counter()
        res, rep = srv.serve(jobs_for(all_idx), workers=args.workers, realtime=wl["realtime"], order=args.order,
                             window=args.window, pin=args.pin, sync_yield="block" if args.sync == "block" else True,
                             precision=wl["precision"])
        walls.append(time.perf_counter() - t)
        served.append((res, rep))
        log(f"[bench] step {s}: {rep['pairs']} pairs in {walls[-1]:.1f}s, workers {rep['workers']}, "
            f"peak concurrency {rep['peak_concurrency']}")
    srv.ktime_enable(False)
    s_cnt, s_ms, s_bytes = srv.ktime_harvest()

    pairs = sum(rep["pairs"] for _, rep in served)
    wall = sum(walls)
    lat = [r["t_done"] - r["arrival"] for res, _ in served for r in res]
    # FP64 residual of every pair: ||A x - l B x|| / (||A|| + |l| ||B||) <= tol
    worst = 0.0
    npairs_checked = 0
    bad = 0
    for res, _ in served:
        for i, r in enumerate(res):
            for lam, pr in zip(r["eigenvalues"], r["pair_residuals"]):
                v = pr / (norms[i] + abs(lam) * 1.0)
                worst = max(worst, v)
                npairs_checked += 1
                bad += v > 1e-10
            bad += sum(r["residuals"] >= 1e-10)
    statuses = [r["status"] for res, _ in served for r in res]
    log(f"[bench] served: {pairs} pairs / {wall:.1f}s; worst normalised residual {worst:.2e}; "
        f"bad {bad}; statuses {sorted(set(statuses))}")

    # ---- plain: one solve at a time, FP64, stratified by (n, k) class ----
    rng = np.random.default_rng(args.seed + 1)
    classes = {}
    for i, r in enumerate(reqs):
        # small solves: n is continuous; eight log-spaced n classes
        key = (r["n"], r["k"]) if args.workload == "serving" else (int(np.log2(r["n"] / 4096) * 8 / 3), r["k"])
        classes.setdefault(key, []).append(i)
    sample, cls_of = [], {}
    for key in sorted(classes):
        members = classes[key]
        pick = rng.choice(len(members), size=min(args.plain_per_class, len(members)), replace=False)
        for j in sorted(pick.tolist()):
            sample.append(members[j])
            cls_of[members[j]] = key
    srv.ktime_enable(True)
    t = time.perf_counter()
    pres, prep = srv.serve(jobs_for(sample, False), workers=1, realtime=False, precision=wl["precision"])
    t_plain = time.perf_counter() - t
    srv.ktime_enable(False)
    kcnt, kms, kbytes = srv.ktime_harvest()
    per_cls = {}
    for j, i in enumerate(sample):
        r = pres[j]
        dt = r["t_done"] - r["t_start"]
        per_cls.setdefault(cls_of[i], []).append((r["nconv"], dt))
        log(f"[bench]   plain n={reqs[i]['n']} k={reqs[i]['k']}: {r['nconv']} pairs, {r['iters']} iters, "
            f"{dt:.2f}s (qz {r['qz_seconds']:.2f}s), status {r['status']}")
    est_pairs = sum(len(classes[c]) * np.mean([p for p, _ in v]) for c, v in per_cls.items())
    est_secs = sum(len(classes[c]) * np.mean([d for _, d in v]) for c, v in per_cls.items())
    plain_rate = est_pairs / est_secs if est_secs else None
    plain_pairs = prep["pairs"]
    diff = None
    if served:
        # same eigenvalues alone and in the mix
        mix = served[0][0]
        diff = 0.0
        for j, i in enumerate(sample):
            a, b = np.sort_complex(pres[j]["eigenvalues"]), np.sort_complex(mix[i]["eigenvalues"])
            if a.size and a.size == b.size:
                diff = max(diff, float(np.max(np.abs(a - b) / np.abs(a))))
        # where served time goes: the sample's solve and QZ seconds alone vs in
        # the mix, and the mean number of solves in flight
        s_alone = sum(pres[j]["t_done"] - pres[j]["t_start"] for j in range(len(sample)))
        s_mix = sum(mix[i]["t_done"] - mix[i]["t_start"] for i in sample)
        q_alone = sum(pres[j]["qz_seconds"] for j in range(len(sample)))
        q_mix = sum(mix[i]["qz_seconds"] for i in sample)
        busy = sum(r["t_done"] - r["t_start"] for r in mix)
        log(f"[bench] sample solve seconds alone {s_alone:.1f} / in the mix {s_mix:.1f} "
            f"(qz {q_alone:.1f} / {q_mix:.1f}); mean solves in flight {busy / walls[0]:.1f}; "
            f"engine stream {sum(r['stream_bytes'] for r in mix) / walls[0] / 1e9:.0f} GB/s over the step")
    log(f"[bench] plain: {len(sample)} solves in {len(per_cls)} classes, {plain_pairs} pairs / {t_plain:.1f}s; "
        f"class-weighted estimate for the workload {plain_rate:.2f} pairs/s; "
        f"max rel eig diff plain vs served {diff}")

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
        if s_cnt:
            roof["served_region_GBs"] = round(s_bytes / (s_ms * 1e-3) / 1e9, 1)
            roof["served_region_avg_us"] = round(s_ms * 1e3 / s_cnt, 2)

    cpu = None
    if not args.no_cpu_baseline and not args.plain_only:
        thr = args.cpu_threads or min(16, os.cpu_count() or 1)
        try:
            cpu = cpu_baseline(reqs, args.cpu_seconds, thr, wl)
            log(f"[bench] cpu baseline: {cpu['value']} eigenpairs/s ({cpu['kind']}, {cpu['cores']} threads)")
        except Exception as e:  # noqa: BLE001
            log(f"[bench] cpu baseline unavailable: {e}")

    out = {
        "metric": "eigenpairs/s @tol1e-10 (served vs plain, 1 GPU); p50/p99 solve latency",
        "value": round(served_rate, 3) if served_rate else None,
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 1) if served else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if args.workload == "serving" else "c128 (outer loop), c64 (inner solve)",
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
                   "host_wait": args.sync, "workers_pinned": bool(args.pin)},
        "served_eigenpairs_per_s": round(served_rate, 3) if served_rate else None,
        "plain_eigenpairs_per_s": round(plain_rate, 3) if plain_rate else None,
        "served_over_plain": round(served_rate / plain_rate, 2) if served_rate and plain_rate else None,
        "plain_sample_requests": len(sample),
        "plain_estimate": "per class mean of its sampled solves alone, weighted by class counts",
        "latency_p50_s": round(pct(lat, 50), 3) if lat else None,
        "latency_p99_s": round(pct(lat, 99), 3) if lat else None,
        "pairs": pairs,
        "max_normalised_residual": worst,
        "pairs_over_tol": int(bad),
        "max_rel_eig_diff_alone_vs_mix": diff,
        "setup_seconds": round(t_setup, 1),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
