/*
 * bbjd_serve.h -- continuous-batching serving loop for independent solves.
 *
 * No counterpart in the reference: there, each shift sigma of an
 * eigenspectrum sweep is an independent solve on its own MPI rank
 * (reference README.md:59-83, include/bbjd.h:14-37).  Here the solves are
 * requests served on one GPU: each worker owns a host thread, a HIP stream
 * and a resident workspace sized for the largest request; requests are
 * admitted in arrival order as workers and HBM free up, so the host QZ of
 * one request overlaps the device work of the others.
 */
#ifndef BBJD_SERVE_H
#define BBJD_SERVE_H

#include "bbjd_ops.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* problem: a resident operator, or a host pencil the server uploads and
     * factors at admission (and frees at completion) */
    bbjd_op                *op;
    const bbjd_pencil_desc *pencil;
    int                     pc;           /* bbjd_pc_kind for an on-demand op  */
    bbjd_zomplex            sigma;
    bbjd_int                nev, maxsub, minsub, maxiter, inner_maxiter;
    double                  tol, inner_tol;
    double                  arrival;      /* seconds after the start of the run */
    /* outputs (caller-allocated arrays of nev entries; eigenvectors n x nev
     * column-major or NULL) */
    bbjd_zomplex           *eigenvalues;
    double                 *residuals;       /* reference criterion           */
    double                 *pair_residuals;  /* ||A x - lambda B x||, ||x|| = 1 */
    bbjd_zomplex           *eigenvectors;
    bbjd_int                nconv, iters;
    int                     status;
    int                     worker;
    double                  t_admit, t_start, t_done;   /* seconds after start */
    double                  qz_seconds;
    double                  stream_bytes;  /* algorithmic HBM bytes of the engine's
                                              n-sized kernels (bbjd_stats) */
    const bbjd_zomplex     *v0;            /* host start vector (n) or NULL:
                                              bbjd_params.v0 of the solve    */
    /* arrival waves: a request's arrival counts from the moment the last
     * request of every earlier wave was admitted (wave 0: from the start of
     * the run), so successive batches of a stream arrive as the backlog of
     * the one before drains instead of after it has finished.  Waves are
     * admitted in increasing order. */
    int                     wave;
    double                  t_arrive;      /* output: effective arrival, seconds
                                              after the start of the run     */
} bbjd_request;

typedef struct {
    int    workers;       /* concurrent solves (0: 16)                       */
    size_t hbm_budget;    /* bytes for workspaces + on-demand operators
                             (0: 90% of the free device memory at start)     */
    int    precision;     /* bbjd_precision of every request                 */
    int    realtime;      /* 1: admit a request no earlier than its arrival  */
    int    sync_yield;    /* 1: host threads yield while waiting on the GPU,
                             2: the workers block (sleep until the GPU
                             signals: bbjd_gpu_options.blocking_sync of
                             every workspace)                                 */
    int    graphs;        /* bbjd_gpu_options.graphs of every request        */
    int    order;         /* which arrived request a free worker takes:
                             BBJD_ORDER_ARRIVAL (first come, first served),
                             BBJD_ORDER_LARGEST (largest estimated work:
                             nev, then n) or BBJD_ORDER_WINDOW (the largest
                             of the `window` longest-waiting ones)           */
    int    pin;           /* 1: worker i runs on the i-th CPU the process may
                             use                                             */
    int    window;        /* BBJD_ORDER_WINDOW: how many of the longest-
                             waiting requests are compared (default 8); a
                             request is overtaken by at most window - 1
                             later arrivals                                  */
    int    qz_mode;       /* bbjd_gpu_options.qz_mode of every request
                             (BBJD_QZ_RITZ by default)                       */
    int    inner_form;    /* bbjd_gpu_options.inner_form of every request
                             (BBJD_INNER_REFERENCE by default)               */
    int    progress;      /* 1: a line on stderr as each arrival wave's last
                             request completes                               */
    /* grouped inner solves (BBJD_PREC_MIXED): > 0 runs the workers' inner
     * solves through a bbjd_batcher of inner_batchers threads (0: 2) taking
     * up to inner_batch solves per batch and waiting up to inner_window_us
     * for more (bbjd_gpu_options.inner_batcher); 0: every worker launches
     * its own kernels                                                     */
    int    inner_batch;
    int    inner_batchers;
    double inner_window_us;
} bbjd_serve_config;

enum { BBJD_ORDER_ARRIVAL = 0, BBJD_ORDER_LARGEST = 1, BBJD_ORDER_WINDOW = 2 };

typedef struct {
    double   wall_seconds;
    bbjd_int pairs;         /* converged pairs, all requests                 */
    bbjd_int failed;        /* requests with a negative status               */
    int      workers;       /* workers actually run (after the HBM budget)   */
    size_t   workspace_bytes;
    size_t   peak_bytes;    /* workspaces + on-demand operators at the peak  */
    int      peak_concurrency;
} bbjd_serve_report;

void bbjd_serve_config_init(bbjd_serve_config *c);

/* Serves nreq requests; returns 0 when the run itself worked (individual
 * request failures are in their status), -1 on a set-up error.  Same as
 * bbjd_server_create + bbjd_server_run + bbjd_server_destroy. */
int bbjd_serve(bbjd_request *reqs, int nreq, const bbjd_serve_config *cfg, bbjd_serve_report *rep);

/* A resident server: its worker threads, HIP streams and workspaces stay
 * allocated across runs, so batches of requests arriving one after another
 * do not pay the workspace set-up (and do not allocate or free device
 * memory) per batch.
 *
 * bbjd_server_create sizes the workspaces for the largest of the nreq
 * requests given (n, nev, maxsub, inner_maxiter; only these fields are
 * read) and takes workers, hbm_budget, precision, graphs, pin and sync_yield
 * from cfg; NULL on failure.  bbjd_server_run serves a batch like
 * bbjd_serve, using cfg's realtime, order and window; a request larger
 * than the capacity ends with status BBJD_ERR_INVALID.  Arrival times are
 * relative to the start of each run.  One run at a time per server. */
typedef struct bbjd_server bbjd_server;
bbjd_server *bbjd_server_create(const bbjd_request *reqs, int nreq, const bbjd_serve_config *cfg);
int          bbjd_server_run(bbjd_server *s, bbjd_request *reqs, int nreq, const bbjd_serve_config *cfg,
                             bbjd_serve_report *rep);
void         bbjd_server_destroy(bbjd_server *s);
int          bbjd_server_workers(const bbjd_server *s);           /* workers with a workspace */
/* the server's inner-solve batcher's counters (zeros without one) */
void         bbjd_server_batch_stats(const bbjd_server *s, bbjd_batcher_stats *st);
size_t       bbjd_server_workspace_bytes(const bbjd_server *s);   /* per worker               */

/* Free device memory now (hipMemGetInfo), for logs. */
size_t bbjd_device_free_bytes(void);

#ifdef __cplusplus
}
#endif
#endif
