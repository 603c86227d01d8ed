/*
 * serve.c -- continuous-batching serving loop (include/bbjd_serve.h).
 *
 * The reference runs the independent solves of an eigenspectrum sweep on
 * separate MPI ranks (reference README.md:59-83).  On one MI355X the solve is
 * bound by the host QZ of its small projected pencil while the GPU idles, so
 * many solves are served at once: W workers, each a host thread with its own
 * HIP stream and a resident workspace sized for the largest request.  The
 * caller's thread admits requests in arrival order whenever a worker is free
 * and the HBM budget holds the request's operator (when the server builds
 * it); a finished request frees its slot immediately and the next one is
 * admitted while the others are still iterating.  Requests never share
 * device memory, so each result is the same alone and in the mix: the
 * engine's reductions depend only on the request's shape.
 *
 * A server (bbjd_server_create) keeps its workers, streams and workspaces
 * resident across runs (bbjd_server_run), so a stream of request batches
 * pays the workspace set-up once; bbjd_serve is create + run + destroy.
 */
#include "../../include/bbjd_serve.h"
#include "engine.h"

#include <complex.h>
#include <math.h>
#include <omp.h>
#include <pthread.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

int openblas_set_num_threads_local(int n);

typedef double complex zcx;

typedef struct {
    bbjd_server    *S;
    int             id;
    pthread_t       th;
    int             started;
    bbjd_workspace *ws;
    int             req;        /* assigned request or -1 */
    size_t          held;       /* bytes charged for its on-demand operator */
    int             ok;
} worker_t;

struct bbjd_server {
    pthread_mutex_t   mu;
    pthread_cond_t    cv_work, cv_free;
    bbjd_request     *reqs;             /* the current run's requests        */
    int               stop, active, peak_active;
    size_t            budget, used, peak, wsb;
    double            t0;
    worker_t         *w;
    int               nw;
    bbjd_params       cap;
    bbjd_gpu_options  opt;
    int               init_failed;
    int               pin;              /* workers pinned to one CPU each    */
    /* the current run's arrival waves: requests and pairs left per wave
     * (index wave - wave_min), reported as each wave completes */
    int              *wave_left;
    int64_t          *wave_pairs;
    int               wave_min, nwaves, progress;
    bbjd_batcher     *batcher;          /* grouped inner solves, or NULL     */
};

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int bbjd_device_init(int sync_yield)
{
    if (sync_yield == 2) hipSetDeviceFlags(hipDeviceScheduleBlockingSync);
    else if (sync_yield) hipSetDeviceFlags(hipDeviceScheduleYield);
    hipFree(0);
    return 0;
}

size_t bbjd_device_free_bytes(void)
{
    size_t freeb = 0, totb = 0;
    return hipMemGetInfo(&freeb, &totb) == hipSuccess ? freeb : 0;
}

void bbjd_serve_config_init(bbjd_serve_config *c)
{
    memset(c, 0, sizeof(*c));
    c->workers = 16;
    c->precision = BBJD_PREC_FP64;
    c->realtime = 1;
    c->sync_yield = 1;
    c->order = BBJD_ORDER_ARRIVAL;
    c->window = 8;
    c->inner_window_us = 100;
}

static bbjd_int req_n(const bbjd_request *q) { return q->op ? bbjd_op_size(q->op) : (q->pencil ? q->pencil->n : 0); }

static bbjd_int req_maxsub(const bbjd_request *q)
{
    bbjd_int ms = 3 * q->nev + 10;
    if (ms < 40) ms = 40;
    const bbjd_int n = req_n(q);
    return q->maxsub > 0 ? q->maxsub : (ms < n ? ms : n);
}

static void run_request(bbjd_server *S, worker_t *w, bbjd_request *q)
{
    q->t_start = now_s() - S->t0;
    q->worker = w->id;
    bbjd_op *op = q->op;
    int own = 0;
    hipStream_t st = (hipStream_t)bbjd_workspace_stream(w->ws);
    if (!op) {
        op = bbjd_op_create(q->pencil, q->sigma, (bbjd_pc_kind)q->pc, S->opt.precision == BBJD_PREC_MIXED, st);
        own = 1;
        if (!op) { q->status = BBJD_ERR_ALLOC; q->t_done = now_s() - S->t0; return; }
    }
    bbjd_params p;
    bbjd_params_init(&p);
    p.n = bbjd_op_size(op);
    p.nev = q->nev;
    p.sigma = q->sigma;
    p.maxsub = q->maxsub;
    p.minsub = q->minsub;
    p.maxiter = q->maxiter > 0 ? q->maxiter : 200;
    p.tol = q->tol > 0 ? q->tol : 1e-10;
    p.inner_maxiter = q->inner_maxiter;
    p.inner_tol = q->inner_tol > 0 ? q->inner_tol : 1e-2;
    p.matvec_A = bbjd_op_matvec_A;
    p.matvec_B = bbjd_op_matvec_B;
    p.precond = bbjd_op_precond;
    p.op_data = op;
    p.v0 = q->v0;
    bbjd_result R;
    memset(&R, 0, sizeof(R));
    q->status = bbjd_solve_in(&p, w->ws, &R);
    q->nconv = R.nconv;
    q->iters = R.iters;
    if (R.nconv > 0) {
        if (q->eigenvalues) memcpy(q->eigenvalues, R.eigenvalues, (size_t)R.nconv * sizeof(zcx));
        if (q->residuals) memcpy(q->residuals, R.residuals, (size_t)R.nconv * sizeof(double));
        if (q->eigenvectors) memcpy(q->eigenvectors, R.eigenvectors, (size_t)R.nconv * p.n * sizeof(zcx));
        if (q->pair_residuals) bbjd_workspace_pair_residuals(w->ws, q->pair_residuals);
    }
    bbjd_stats stt;
    if (bbjd_workspace_stats(w->ws, &stt) == 0) {
        q->qz_seconds = stt.qz_seconds;
        q->stream_bytes = stt.stream_bytes;
    }
    bbjd_result_free(&R);
    if (own) bbjd_op_free(op);
    q->t_done = now_s() - S->t0;
}

static void *worker_main(void *arg)
{
    worker_t *w = (worker_t *)arg;
    bbjd_server *S = w->S;
    /* one LAPACK thread and one OpenMP thread per worker: the workers are
     * the parallelism */
    openblas_set_num_threads_local(1);
    omp_set_num_threads(1);
    if (S->pin) {
        /* worker i on the i-th CPU this process may use */
        cpu_set_t allowed, one;
        if (sched_getaffinity(0, sizeof(allowed), &allowed) == 0) {
            const int cnt = CPU_COUNT(&allowed);
            int want = cnt > 0 ? w->id % cnt : 0, seen = 0;
            for (int c = 0; c < CPU_SETSIZE && cnt > 0; c++)
                if (CPU_ISSET(c, &allowed) && seen++ == want) {
                    CPU_ZERO(&one);
                    CPU_SET(c, &one);
                    (void)pthread_setaffinity_np(pthread_self(), sizeof(one), &one);
                    break;
                }
        }
    }
    w->ws = bbjd_workspace_alloc_gpu(&S->cap, &S->opt);
    pthread_mutex_lock(&S->mu);
    w->ok = w->ws != NULL;
    if (!w->ok) S->init_failed++;
    pthread_cond_broadcast(&S->cv_free);
    for (;;) {
        while (w->req < 0 && !S->stop) pthread_cond_wait(&S->cv_work, &S->mu);
        if (w->req < 0 && S->stop) break;
        const int r = w->req;
        bbjd_request *q = &S->reqs[r];
        pthread_mutex_unlock(&S->mu);
        run_request(S, w, q);
        pthread_mutex_lock(&S->mu);
        if (S->wave_left) {
            const int wi = q->wave - S->wave_min;
            S->wave_pairs[wi] += q->nconv;
            if (--S->wave_left[wi] == 0 && S->progress) {
                fprintf(stderr, "[serve] wave %d done at %.1f s: %lld pairs\n", q->wave, q->t_done,
                        (long long)S->wave_pairs[wi]);
                fflush(stderr);
            }
        }
        S->used -= w->held;
        w->held = 0;
        w->req = -1;
        S->active--;
        pthread_cond_broadcast(&S->cv_free);
    }
    pthread_mutex_unlock(&S->mu);
    if (w->ws) bbjd_workspace_free(w->ws);
    w->ws = NULL;
    return NULL;
}

static int cmp_arrival(const void *a, const void *b, void *ctx)
{
    const bbjd_request *R = (const bbjd_request *)ctx;
    const int wa = R[*(const int *)a].wave, wb = R[*(const int *)b].wave;
    if (wa != wb) return wa < wb ? -1 : 1;
    const double x = R[*(const int *)a].arrival, y = R[*(const int *)b].arrival;
    return x < y ? -1 : x > y ? 1 : (*(const int *)a - *(const int *)b);
}

void bbjd_server_destroy(bbjd_server *S)
{
    if (!S) return;
    if (S->w) {
        pthread_mutex_lock(&S->mu);
        S->stop = 1;
        pthread_cond_broadcast(&S->cv_work);
        pthread_mutex_unlock(&S->mu);
        for (int i = 0; i < S->nw; i++)
            if (S->w[i].started) pthread_join(S->w[i].th, NULL);
        free(S->w);
    }
    bbjd_batcher_destroy(S->batcher);
    pthread_mutex_destroy(&S->mu);
    pthread_cond_destroy(&S->cv_work);
    pthread_cond_destroy(&S->cv_free);
    free(S);
}

bbjd_server *bbjd_server_create(const bbjd_request *reqs, int nreq, const bbjd_serve_config *cfg)
{
    if (!reqs || nreq <= 0 || !cfg) return NULL;
    /* the process's wait (ignored once the device is up); blocking waits are
     * the workspaces' own (bbjd_gpu_options.blocking_sync), so other solves
     * of the process keep theirs */
    if (cfg->sync_yield) hipSetDeviceFlags(hipDeviceScheduleYield);
    bbjd_server *S = calloc(1, sizeof(*S));
    if (!S) return NULL;
    pthread_mutex_init(&S->mu, NULL);
    pthread_cond_init(&S->cv_work, NULL);
    pthread_cond_init(&S->cv_free, NULL);
    S->pin = cfg->pin;
    /* capacity of the resident workspaces: the largest of the requests */
    bbjd_params_init(&S->cap);
    S->cap.n = 1; S->cap.nev = 1; S->cap.maxsub = 2; S->cap.inner_maxiter = 0;
    for (int i = 0; i < nreq; i++) {
        const bbjd_request *q = &reqs[i];
        const bbjd_int n = req_n(q);
        if (n <= 0 || q->nev <= 0) { bbjd_server_destroy(S); return NULL; }
        if (n > S->cap.n) S->cap.n = n;
        if (q->nev > S->cap.nev) S->cap.nev = q->nev;
        const bbjd_int ms = req_maxsub(q);
        if (ms > S->cap.maxsub) S->cap.maxsub = ms;
        if (q->inner_maxiter > S->cap.inner_maxiter) S->cap.inner_maxiter = q->inner_maxiter;
    }
    bbjd_gpu_options_init(&S->opt);
    S->opt.callback_space = BBJD_CALLBACK_DEVICE;
    S->opt.precision = cfg->precision;
    S->opt.precond2 = bbjd_op_precond2;
    S->opt.matvec_AB = bbjd_op_matvec_AB;
    S->opt.matvec_ABr = bbjd_op_matvec_ABr;
    S->opt.graphs = cfg->graphs;
    S->opt.qz_mode = cfg->qz_mode;
    S->opt.inner_form = cfg->inner_form;
    S->opt.blocking_sync = cfg->sync_yield == 2;
    if (cfg->precision == BBJD_PREC_MIXED) {
        S->opt.matvec_A_c64 = bbjd_op_matvec_A_c64;
        S->opt.matvec_B_c64 = bbjd_op_matvec_B_c64;
        S->opt.precond_c64 = bbjd_op_precond_c64;
        S->opt.matvec_AB_c64 = bbjd_op_matvec_AB_c64;
        if (cfg->inner_batch > 0 && S->cap.inner_maxiter > 0) {
            S->batcher = bbjd_batcher_create(cfg->inner_batchers > 0 ? cfg->inner_batchers : 2, cfg->inner_batch,
                                             cfg->inner_window_us, S->cap.inner_maxiter);
            if (!S->batcher) { bbjd_server_destroy(S); return NULL; }
            S->opt.inner_batcher = S->batcher;
        }
    }
    S->wsb = bbjd_workspace_bytes(&S->cap, &S->opt);
    size_t freeb = 0, totb = 0;
    if (hipMemGetInfo(&freeb, &totb) != hipSuccess) { bbjd_server_destroy(S); return NULL; }
    S->budget = cfg->hbm_budget ? cfg->hbm_budget : (size_t)(0.9 * (double)freeb);
    int nw = cfg->workers > 0 ? cfg->workers : 16;
    if (nw > nreq) nw = nreq;
    while (nw > 1 && (size_t)nw * S->wsb > S->budget) nw--;
    if ((size_t)nw * S->wsb > S->budget) { bbjd_server_destroy(S); return NULL; }
    S->w = calloc((size_t)nw, sizeof(worker_t));
    if (!S->w) { bbjd_server_destroy(S); return NULL; }
    S->nw = nw;
    for (int i = 0; i < nw; i++) {
        S->w[i].S = S; S->w[i].id = i; S->w[i].req = -1;
        if (pthread_create(&S->w[i].th, NULL, worker_main, &S->w[i]) == 0) S->w[i].started = 1;
        else {
            pthread_mutex_lock(&S->mu);
            S->init_failed++;
            pthread_mutex_unlock(&S->mu);
        }
    }
    /* wait for the workspaces */
    pthread_mutex_lock(&S->mu);
    for (;;) {
        int up = 0;
        for (int i = 0; i < nw; i++) up += S->w[i].ok;
        if (up + S->init_failed >= nw) break;
        pthread_cond_wait(&S->cv_free, &S->mu);
    }
    const int failed_init = S->init_failed;
    pthread_mutex_unlock(&S->mu);
    if (failed_init == nw) { bbjd_server_destroy(S); return NULL; }
    return S;
}

void bbjd_server_batch_stats(const bbjd_server *S, bbjd_batcher_stats *st)
{
    memset(st, 0, sizeof(*st));
    if (S && S->batcher) bbjd_batcher_get_stats(S->batcher, st);
}

int bbjd_server_workers(const bbjd_server *S)
{
    if (!S) return 0;
    int up = 0;
    for (int i = 0; i < S->nw; i++) up += S->w[i].ok;
    return up;
}

size_t bbjd_server_workspace_bytes(const bbjd_server *S) { return S ? S->wsb : 0; }

int bbjd_server_run(bbjd_server *S, bbjd_request *reqs, int nreq, const bbjd_serve_config *cfg,
                    bbjd_serve_report *rep)
{
    if (!S || !reqs || nreq <= 0 || !cfg || !rep) return -1;
    memset(rep, 0, sizeof(*rep));
    /* every request must fit the resident workspaces */
    for (int i = 0; i < nreq; i++) {
        bbjd_request *q = &reqs[i];
        q->status = 0; q->nconv = 0; q->iters = 0; q->worker = -1;
        q->t_admit = q->t_start = q->t_done = q->t_arrive = 0;
        const bbjd_int n = req_n(q);
        if (n <= 0 || q->nev <= 0) return -1;
    }
    const int nw = S->nw;
    int *order = malloc((size_t)nreq * sizeof(int));
    int *pend = malloc((size_t)nreq * sizeof(int));
    if (!order || !pend) { free(order); free(pend); return -1; }
    for (int i = 0; i < nreq; i++) order[i] = i;
    qsort_r(order, (size_t)nreq, sizeof(int), cmp_arrival, reqs);

    /* per-wave counts for the completion reports */
    int wmin = reqs[0].wave, wmax = reqs[0].wave;
    for (int i = 1; i < nreq; i++) {
        if (reqs[i].wave < wmin) wmin = reqs[i].wave;
        if (reqs[i].wave > wmax) wmax = reqs[i].wave;
    }
    int *wave_left = NULL;
    int64_t *wave_pairs = NULL;
    const long nwv = (long)wmax - (long)wmin + 1;
    if (nwv <= 65536) {
        wave_left = calloc((size_t)nwv, sizeof(int));
        wave_pairs = calloc((size_t)nwv, sizeof(int64_t));
        if (!wave_left || !wave_pairs) { free(wave_left); free(wave_pairs); wave_left = NULL; wave_pairs = NULL; }
        else for (int i = 0; i < nreq; i++) wave_left[reqs[i].wave - wmin]++;
    }
    pthread_mutex_lock(&S->mu);
    S->wave_left = wave_left;
    S->wave_pairs = wave_pairs;
    S->wave_min = wmin;
    S->nwaves = (int)nwv;
    S->progress = cfg->progress;
    S->reqs = reqs;
    S->used = (size_t)nw * S->wsb;
    S->peak = S->used;
    S->active = 0;
    S->peak_active = 0;
    S->t0 = now_s();     /* the run starts with every workspace resident */
    pthread_mutex_unlock(&S->mu);
    rep->workers = bbjd_server_workers(S);
    rep->workspace_bytes = S->wsb;

    int npend = 0, next = 0;
    /* the wave being admitted and the time its arrivals count from */
    int wave = reqs[order[0]].wave;
    double base = 0;
    for (int admitted = 0; admitted < nreq; admitted++) {
        /* a free worker first: the choice among waiting requests is made
         * when one can start */
        pthread_mutex_lock(&S->mu);
        for (;;) {
            int free_slot = 0;
            for (int i = 0; i < nw && !free_slot; i++) free_slot = S->w[i].ok && S->w[i].req < 0;
            if (free_slot) break;
            pthread_cond_wait(&S->cv_free, &S->mu);
        }
        pthread_mutex_unlock(&S->mu);
        /* everything that has arrived by now; at least the next arrival */
        if (npend == 0 && cfg->realtime) {
            const double wait = base + reqs[order[next]].arrival - (now_s() - S->t0);
            if (wait > 0) {
                struct timespec ts = {(time_t)wait, (long)((wait - (double)(time_t)wait) * 1e9)};
                nanosleep(&ts, NULL);
            }
        }
        const double now = now_s() - S->t0;
        while (next < nreq && reqs[order[next]].wave == wave &&
               (!cfg->realtime || base + reqs[order[next]].arrival <= now || npend == 0)) {
            reqs[order[next]].t_arrive = base + reqs[order[next]].arrival;
            pend[npend++] = order[next++];
        }
        int pi = 0;
        int cand = 1;       /* the waiting requests compared, oldest first */
        if (cfg->order == BBJD_ORDER_LARGEST) cand = npend;
        else if (cfg->order == BBJD_ORDER_WINDOW) cand = cfg->window < 1 ? 1 : (cfg->window < npend ? cfg->window : npend);
        if (cand > 1)
            for (int i = 1; i < cand; i++) {
                const bbjd_request *a = &reqs[pend[i]], *b = &reqs[pend[pi]];
                if (a->nev > b->nev || (a->nev == b->nev && req_n(a) > req_n(b))) pi = i;
            }
        const int r = pend[pi];
        memmove(pend + pi, pend + pi + 1, (size_t)(npend - pi - 1) * sizeof(int));
        npend--;
        bbjd_request *q = &reqs[r];
        const size_t need = q->op ? 0 : bbjd_op_bytes_for(q->pencil, q->pc, S->opt.precision == BBJD_PREC_MIXED);
        pthread_mutex_lock(&S->mu);
        int slot = -1;
        for (;;) {
            for (int i = 0; i < nw; i++)
                if (S->w[i].ok && S->w[i].req < 0) { slot = i; break; }
            if (slot >= 0 && (S->used + need <= S->budget || S->active == 0)) break;
            slot = -1;
            pthread_cond_wait(&S->cv_free, &S->mu);
        }
        S->used += need;
        if (S->used > S->peak) S->peak = S->used;
        S->w[slot].held = need;
        S->w[slot].req = r;
        S->active++;
        if (S->active > S->peak_active) S->peak_active = S->active;
        q->t_admit = now_s() - S->t0;
        pthread_cond_broadcast(&S->cv_work);
        pthread_mutex_unlock(&S->mu);
        /* the wave's last request is admitted: the next wave's clock starts */
        if (npend == 0 && next < nreq && reqs[order[next]].wave != wave) {
            wave = reqs[order[next]].wave;
            base = q->t_admit;
        }
    }
    free(pend);
    pthread_mutex_lock(&S->mu);
    while (S->active > 0) pthread_cond_wait(&S->cv_free, &S->mu);
    rep->wall_seconds = now_s() - S->t0;
    rep->peak_bytes = S->peak;
    rep->peak_concurrency = S->peak_active;
    S->reqs = NULL;
    S->wave_left = NULL;
    S->wave_pairs = NULL;
    pthread_mutex_unlock(&S->mu);
    free(wave_left);
    free(wave_pairs);
    for (int i = 0; i < nreq; i++) {
        rep->pairs += reqs[i].nconv;
        if (reqs[i].status < 0) rep->failed++;
    }
    free(order);
    return 0;
}

int bbjd_serve(bbjd_request *reqs, int nreq, const bbjd_serve_config *cfg, bbjd_serve_report *rep)
{
    if (!reqs || nreq <= 0 || !cfg || !rep) return -1;
    memset(rep, 0, sizeof(*rep));
    bbjd_server *S = bbjd_server_create(reqs, nreq, cfg);
    if (!S) return -1;
    const int rc = bbjd_server_run(S, reqs, nreq, cfg, rep);
    bbjd_server_destroy(S);
    return rc;
}
