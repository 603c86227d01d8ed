/*
 * jd_engine.c -- JDQZ outer iteration on one MI355X.
 *
 * Host driver in C11 over the HIP kernels of kernels.hip.  It implements
 * the reference C API (include/bbjd.h) with the algorithm of reference
 * src/jd_core.c:743-1120 -- same steps, thresholds (DGKS 0.7071, breakdown
 * 1e-14, guards 1e-300, lock matching |lambda - lambda_c| < 100 tol), same
 * start vector (glibc srand(42)/rand() sequence, reproduced thread-safely
 * with random_r) and same host LAPACK for the projected QZ -- but with
 * V, W, AV, BV, Q_right, Q_left, the GMRES basis and the n-vectors in HBM.
 *
 * Per outer iteration the host/device traffic is: the k Ritz coefficients up,
 * two norms down, the new pencil row/column down (4k+2 complex), i.e. the
 * host never touches an n-sized array unless callbacks live on the host.
 */
#include "../../include/bbjd.h"
#include "../../include/bbjd_gpu.h"
#include "../../include/bbjd_ops.h"
#include "engine.h"
#include "lapack_host.h"
#include "qz_host.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <pthread.h>
#include <string.h>
#include <time.h>

typedef double complex zcx;
typedef struct { float re, im; } c64;

/* Arnoldi steps of the inner solve the host may have in flight at once */
#define GM_LOOK_MAX 8
static int g_gmres_look = 4;
void bbjd_set_gmres_lookahead(int steps) { g_gmres_look = steps < 1 ? 1 : steps > GM_LOOK_MAX ? GM_LOOK_MAX : steps; }

struct bbjd_dev {
    hipStream_t      stream;
    int              own_stream;
    bbjd_gpu_options opt;
    int              ims;
    void            *dblock;           /* the single device allocation     */
    size_t           dbytes;
    bbjd_red         rd;
    void            *zr;               /* restart rotation, maxsub^2       */
    void            *pencil;           /* 4 (maxsub + 1) complex            */
    void            *gram;             /* 2 maxsub^2 complex                */
    /* mixed precision */
    void            *gvc;              /* complex64 basis ld x (ims + 1)   */
    void            *qrc, *qlc;        /* complex64 projectors             */
    void            *tc[3];            /* complex64 n-vectors               */
    /* projected-preconditioner inner solve (BBJD_INNER_PROJECTED): Yh =
     * M^-1 Q_left (complex64, nev + 1 columns), H^-1 (complex64, ndefl^2),
     * staging of H's new entries; host: H (complex128, ld nev + 1), pinned
     * copies; pp_valid = leading columns of Yh and H that hold locked
     * columns (the last one of a call is the current Ritz pair's, redone) */
    void            *yc, *hinv, *hnew;
    zcx             *hh;
    c64             *hpp;
    bbjd_int         pp_ld, pp_valid;
    /* host side */
    double          *hscal;            /* pinned: SL_COUNT doubles + flags  */
    zcx             *hsmall;           /* pinned: small vector staging      */
    zcx             *hx, *hy;          /* pinned: host-callback staging     */
    zcx             *al, *be, *yv, *xv;
    zcx             *zrh, *hcol, *hgram;  /* pinned: rotation, pencil, Gram */
    bbjd_qz          qz;
    struct random_data rng;
    int32_t          rng_state[32];
    bbjd_stats       st;
    double          *pair_res;         /* per converged pair, last solve    */
    bbjd_int         nconv_last;
    /* captured inner-solve segments (bbjd_gpu_options.graphs): [0] the
     * GMRES prologue, [1 + j] Arnoldi step j, valid for gkey */
    hipGraphExec_t  *gx;
    int              gx_cap;
    struct gkey { bbjd_int ndefl, n, ld; const void *r, *z, *qr, *ql, *op; int pp; } gkey;
    double          *gco;              /* device: theta_b, -theta_a         */
    double          *hgco;             /* pinned: their staging             */
    /* the projected step on the device (qz_mode BBJD_QZ_DEVICE): MA, MB,
     * the converged values and the outputs in HBM, the problem descriptor
     * and the outputs' staging pinned */
    void            *qzd;
    char            *hqzd;
    int              zr_ready;          /* d->zr holds the restart's rotation  */
    /* inner-solve lookahead: per in-flight Arnoldi step, a pinned slot for
     * its Hessenberg column and scalar slots, and an event after its copies */
    char            *hlook;
    size_t           look_slot;
    hipEvent_t       lev[GM_LOOK_MAX];
    int              nlev;
    /* bbjd_gpu_options.blocking_sync: host waits on the workspace's stream
     * sleep on this event (created with hipEventBlockingSync) instead of
     * the process's default wait */
    hipEvent_t       bev;
};

static __thread hipStream_t tl_cb_stream;

/* Waits for the work enqueued on s (a workspace's stream): through the
 * workspace's blocking event when it has one, else hipStreamSynchronize. */
static hipError_t dsync(const struct bbjd_dev *d, hipStream_t s)
{
    if (d->bev) {
        const hipError_t e = hipEventRecord(d->bev, s);
        return e != hipSuccess ? e : hipEventSynchronize(d->bev);
    }
    return hipStreamSynchronize(s);
}
void *bbjd_callback_stream(void) { return (void *)tl_cb_stream; }

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

#define HIPOK(x) ((x) == hipSuccess)

/* ------------------------------------------------------------------ */
/* allocation                                                          */
/* ------------------------------------------------------------------ */
typedef struct {
    bbjd_int n, ld, maxsub, nev, ims;
    int mixed, proj;
    size_t off[48];
    size_t total;
} layout_t;

/* O_PEN directly precedes O_SCAL: the pencil column and the scalar slots
 * come back to the host in one copy */
enum { O_V, O_W, O_AV, O_BV, O_QR, O_QL, O_GV, O_T1, O_T2, O_T3, O_T4, O_HW, O_PART, O_PNORM, O_H, O_H2,
       O_HSUM, O_PEN, O_SCAL, O_FLAG, O_CNT, O_COEF, O_COEF2, O_ZR, O_GRAM, O_GVC, O_QRC, O_QLC, O_TC0, O_TC1,
       O_TC2, O_GCO, O_QZD, O_YC, O_HINV, O_HNEW, O_END };
_Static_assert(O_END <= 48, "layout_t.off holds every sub-buffer");

static bbjd_int auto_maxsub(const bbjd_params *p)
{
    bbjd_int ms = 3 * p->nev + 10;      /* reference src/jd_core.c:95-98 */
    if (ms < 40) ms = 40;
    return p->maxsub > 0 ? p->maxsub : (ms < p->n ? ms : p->n);
}

/* device region of the projected step: MA, MB (maxsub^2 each), the
 * converged values (nev + 1), alpha, beta, x (maxsub each), 4 ints and the
 * problem descriptor */
static size_t qzd_bytes(size_t ms, size_t nev)
{
    return (2 * ms * ms + nev + 1 + 3 * ms) * 16 + 512 + (sizeof(bbjd_qz_problem) + 255) / 256 * 256;
}
/* pinned staging of the same: descriptor, info, alpha, beta, x */
static size_t hqzd_bytes(size_t ms) { return 512 + (sizeof(bbjd_qz_problem) + 255) / 256 * 256 + 3 * ms * 16; }

static void make_layout(layout_t *L, const bbjd_params *p, const bbjd_gpu_options *o)
{
    memset(L, 0, sizeof(*L));
    L->n = p->n;
    L->ld = (p->n + 63) / 64 * 64;
    L->maxsub = auto_maxsub(p);
    L->nev = p->nev;
    L->ims = p->inner_maxiter > 0 ? p->inner_maxiter : 1;
    L->mixed = o && o->precision == BBJD_PREC_MIXED;
    L->proj = L->mixed && o->inner_form == BBJD_INNER_PROJECTED && p->inner_maxiter > 0;
    const size_t Z = 16, C = 8, ld = (size_t)L->ld, ms = (size_t)L->maxsub;
    size_t sz[O_END];
    memset(sz, 0, sizeof(sz));
    sz[O_V] = sz[O_W] = sz[O_AV] = sz[O_BV] = ld * ms * Z;
    sz[O_QR] = sz[O_QL] = ld * (size_t)(L->nev + 1) * Z;
    sz[O_GV] = L->mixed ? 0 : ld * (size_t)(L->ims + 1) * Z;
    sz[O_T1] = sz[O_T2] = sz[O_T3] = sz[O_T4] = ld * Z;
    sz[O_HW] = (2 * ms + (size_t)L->nev + 4) * Z;
    int maxk = (int)(ms > (size_t)L->ims + 1 ? ms : (size_t)L->ims + 1);
    if (maxk < L->nev + 1) maxk = (int)L->nev + 1;
    sz[O_PART] = (size_t)bbjd_red_part_entries(L->n, maxk + 2) * Z;
    sz[O_PNORM] = ((size_t)(L->n + BBJD_TILE_MIN_ROWS - 1) / BBJD_TILE_MIN_ROWS + 64) * 8;
    sz[O_H] = sz[O_H2] = sz[O_HSUM] = ((size_t)maxk + 8) * Z;
    sz[O_SCAL] = SL_COUNT * 8;
    sz[O_FLAG] = FL_COUNT * 4;
    sz[O_CNT] = BBJD_CNT_SLOTS * 4;
    sz[O_COEF] = ((size_t)maxk + 8) * Z;
    sz[O_COEF2] = 64 * Z;
    sz[O_ZR] = ms * ms * Z;
    sz[O_PEN] = 4 * (ms + 1) * Z;
    sz[O_GRAM] = 2 * ms * ms * Z;
    sz[O_GCO] = 64;
    sz[O_QZD] = qzd_bytes(ms, (size_t)L->nev);
    if (L->mixed) {
        sz[O_GVC] = ld * (size_t)(L->ims + 1) * C;
        sz[O_QRC] = sz[O_QLC] = ld * (size_t)(L->nev + 1) * C;
        sz[O_TC0] = sz[O_TC1] = sz[O_TC2] = ld * C;
    }
    if (L->proj) {
        /* projected preconditioner: M^-1 Q_left (complex64), H^-1 and the
         * new entries of H */
        const size_t K = (size_t)L->nev + 1;
        sz[O_YC] = ld * K * C;
        sz[O_HINV] = K * K * C;
        sz[O_HNEW] = 2 * K * K * C;
    }
    size_t off = 0;
    for (int i = 0; i < O_END; i++) {
        L->off[i] = off;
        off += (sz[i] + 255) / 256 * 256;
    }
    L->total = off;
}

void bbjd_params_init(bbjd_params *p)
{
    /* reference src/jd_core.c:73-84 */
    memset(p, 0, sizeof(*p));
    p->nev = 6;
    p->sigma = 0.0;
    p->target = BBJD_TARGET_NEAREST;
    p->tol = 1e-10;
    p->maxiter = 200;
    p->inner_maxiter = 0;
    p->inner_tol = 1e-2;
    p->v0 = NULL;
}

void bbjd_gpu_options_init(bbjd_gpu_options *o)
{
    memset(o, 0, sizeof(*o));
    o->callback_space = BBJD_CALLBACK_HOST;
    o->precision = BBJD_PREC_FP64;
}

size_t bbjd_workspace_bytes(const bbjd_params *p, const bbjd_gpu_options *o)
{
    layout_t L;
    make_layout(&L, p, o);
    return L.total;
}

static pthread_rwlock_t g_capture_rw = PTHREAD_RWLOCK_INITIALIZER;
static int g_graph_workspaces;      /* workspaces with graphs on; changed
                                       under the exclusive lock only */
void bbjd_capture_lock(void) { pthread_rwlock_rdlock(&g_capture_rw); }
void bbjd_capture_unlock(void) { pthread_rwlock_unlock(&g_capture_rw); }
int bbjd_devwide_lock(void)
{
    if (__atomic_load_n(&g_graph_workspaces, __ATOMIC_ACQUIRE) == 0) return 0;
    pthread_rwlock_wrlock(&g_capture_rw);
    return 1;
}
void bbjd_devwide_unlock(int held)
{
    if (held) pthread_rwlock_unlock(&g_capture_rw);
}

static void graphs_reset(struct bbjd_dev *d)
{
    for (int i = 0; i < d->gx_cap; i++)
        if (d->gx[i]) { (void)hipGraphExecDestroy(d->gx[i]); d->gx[i] = NULL; }
    memset(&d->gkey, 0, sizeof(d->gkey));
}

/* graphs_reset outside workspace release: destroying executable graphs
 * frees device memory, so it excludes captures on other threads */
static void graphs_drop(struct bbjd_dev *d)
{
    int any = 0;
    for (int i = 0; i < d->gx_cap && !any; i++) any = d->gx[i] != NULL;
    if (!any) { memset(&d->gkey, 0, sizeof(d->gkey)); return; }
    pthread_rwlock_wrlock(&g_capture_rw);
    graphs_reset(d);
    pthread_rwlock_unlock(&g_capture_rw);
}

static void workspace_free_impl(bbjd_workspace *ws)
{
    if (!ws) return;
    struct bbjd_dev *d = ws->dev;
    if (d) {
        if (d->stream) (void)hipStreamSynchronize(d->stream);   /* copies into pinned buffers */
        if (d->gx) { graphs_reset(d); free(d->gx); }
        if (d->hgco) hipHostFree(d->hgco);
        if (d->hqzd) hipHostFree(d->hqzd);
        if (d->hlook) hipHostFree(d->hlook);
        for (int i = 0; i < d->nlev; i++) (void)hipEventDestroy(d->lev[i]);
        if (d->bev) (void)hipEventDestroy(d->bev);
        if (d->dblock) hipFree(d->dblock);
        if (d->hscal) hipHostFree(d->hscal);
        if (d->hsmall) hipHostFree(d->hsmall);
        if (d->hx) hipHostFree(d->hx);
        if (d->hy) hipHostFree(d->hy);
        if (d->own_stream && d->stream) hipStreamDestroy(d->stream);
        if (d->zrh) hipHostFree(d->zrh);
        if (d->hcol) hipHostFree(d->hcol);
        if (d->hgram) hipHostFree(d->hgram);
        if (d->hpp) hipHostFree(d->hpp);
        free(d->hh);
        free(d->al); free(d->be); free(d->yv); free(d->xv); free(d->pair_res);
        bbjd_qz_free(&d->qz);
        free(d);
    }
    free(ws->MA); free(ws->MB); free(ws->SA); free(ws->SB); free(ws->ZA); free(ws->ZB);
    free(ws->gmres_H); free(ws->gmres_cs); free(ws->gmres_sn); free(ws->gmres_g);
    free(ws);
}

/* Workspaces always take the exclusive lock: they change the count of
 * graph-enabled workspaces that gates it. */
void bbjd_workspace_free(bbjd_workspace *ws)
{
    if (!ws) return;
    pthread_rwlock_wrlock(&g_capture_rw);
    if (ws->dev && ws->dev->opt.graphs) __atomic_sub_fetch(&g_graph_workspaces, 1, __ATOMIC_RELEASE);
    workspace_free_impl(ws);
    pthread_rwlock_unlock(&g_capture_rw);
}

static bbjd_workspace *workspace_alloc_impl(const bbjd_params *p, const bbjd_gpu_options *o)
{
    if (!p || p->n <= 0 || p->nev <= 0) return NULL;
    bbjd_gpu_options dflt;
    bbjd_gpu_options_init(&dflt);
    if (!o) o = &dflt;
    layout_t L;
    make_layout(&L, p, o);
    if (L.maxsub > BBJD_MAXK - 2 || L.ims > BBJD_MAXK - 2) return NULL;

    bbjd_workspace *ws = calloc(1, sizeof(*ws));
    if (!ws) return NULL;
    struct bbjd_dev *d = calloc(1, sizeof(*d));
    if (!d) { free(ws); return NULL; }
    ws->dev = d;
    d->opt = *o;
    d->ims = (int)L.ims;
    ws->n = L.n; ws->maxsub = L.maxsub; ws->nev = L.nev; ws->ld = L.ld;
    ws->inner_maxiter = L.ims;

    if (o->stream) d->stream = (hipStream_t)o->stream;
    else {
        if (!HIPOK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking))) goto fail;
        d->own_stream = 1;
    }
    if (!HIPOK(hipMalloc(&d->dblock, L.total))) { d->dblock = NULL; goto fail; }
    d->dbytes = L.total;
    char *b = (char *)d->dblock;
#define AT(i) ((void *)(b + L.off[i]))
    ws->V = AT(O_V); ws->W = AT(O_W); ws->AV = AT(O_AV); ws->BV = AT(O_BV);
    ws->Q_right = AT(O_QR); ws->Q_left = AT(O_QL);
    ws->gmres_V = L.mixed ? AT(O_GVC) : AT(O_GV);
    ws->tmp1 = AT(O_T1); ws->tmp2 = AT(O_T2); ws->tmp3 = AT(O_T3); ws->tmp4 = AT(O_T4);
    ws->h_work = AT(O_HW);
    ws->swap_nm = NULL;
    d->rd.part = AT(O_PART);
    d->rd.pnorm = AT(O_PNORM);
    d->rd.h = AT(O_H); d->rd.h2 = AT(O_H2); d->rd.hsum = AT(O_HSUM);
    d->rd.scal = AT(O_SCAL); d->rd.flag = AT(O_FLAG); d->rd.cnt = AT(O_CNT);
    d->rd.coef = AT(O_COEF); d->rd.coef2 = AT(O_COEF2);
    {
        int maxk = (int)(L.maxsub > L.ims + 1 ? L.maxsub : L.ims + 1);
        if (maxk < L.nev + 1) maxk = (int)L.nev + 1;
        d->rd.part_cap = bbjd_red_part_entries(L.n, maxk + 2);
    }
    d->zr = AT(O_ZR); d->pencil = AT(O_PEN); d->gram = AT(O_GRAM); d->gco = AT(O_GCO); d->qzd = AT(O_QZD);
    if (L.mixed) {
        d->gvc = AT(O_GVC); d->qrc = AT(O_QRC); d->qlc = AT(O_QLC);
        d->tc[0] = AT(O_TC0); d->tc[1] = AT(O_TC1); d->tc[2] = AT(O_TC2);
    }
    if (L.proj) {
        d->yc = AT(O_YC); d->hinv = AT(O_HINV); d->hnew = AT(O_HNEW);
        d->pp_ld = L.nev + 1;
    }
#undef AT
    if (!HIPOK(hipMemsetAsync(d->rd.scal, 0, SL_COUNT * 8, d->stream))) goto fail;
    if (!HIPOK(hipMemsetAsync(d->rd.flag, 0, FL_COUNT * 4, d->stream))) goto fail;
    if (!HIPOK(hipMemsetAsync(d->rd.cnt, 0, BBJD_CNT_SLOTS * 4, d->stream))) goto fail;
    if (!HIPOK(hipHostMalloc((void **)&d->hscal, 256, 0))) { d->hscal = NULL; goto fail; }
    const size_t small = (size_t)(6 * L.maxsub + 2 * L.ims + 64);
    if (!HIPOK(hipHostMalloc((void **)&d->hsmall, small * sizeof(zcx), 0))) { d->hsmall = NULL; goto fail; }
    if (o->callback_space == BBJD_CALLBACK_HOST) {
        if (!HIPOK(hipHostMalloc((void **)&d->hx, (size_t)L.n * sizeof(zcx), 0))) { d->hx = NULL; goto fail; }
        if (!HIPOK(hipHostMalloc((void **)&d->hy, (size_t)L.n * sizeof(zcx), 0))) { d->hy = NULL; goto fail; }
    }
    const size_t ss = (size_t)L.maxsub * L.maxsub;
    ws->MA = calloc(ss, sizeof(zcx)); ws->MB = calloc(ss, sizeof(zcx));
    ws->SA = calloc(ss, sizeof(zcx)); ws->SB = calloc(ss, sizeof(zcx));
    ws->ZA = calloc(ss, sizeof(zcx)); ws->ZB = calloc(ss, sizeof(zcx));
    ws->gmres_H = calloc((size_t)(L.ims + 1) * L.ims, sizeof(zcx));
    ws->gmres_cs = calloc((size_t)L.ims, sizeof(zcx));
    ws->gmres_sn = calloc((size_t)L.ims, sizeof(zcx));
    ws->gmres_g = calloc((size_t)L.ims + 1, sizeof(zcx));
    d->al = malloc((size_t)L.maxsub * sizeof(zcx)); d->be = malloc((size_t)L.maxsub * sizeof(zcx));
    d->yv = malloc((size_t)(L.maxsub + L.ims + 2) * sizeof(zcx));
    d->xv = malloc((size_t)(L.maxsub + 2) * sizeof(zcx));
    if (!HIPOK(hipHostMalloc((void **)&d->zrh, ss * sizeof(zcx), 0))) { d->zrh = NULL; goto fail; }
    /* pencil column/row staging + the scalar slots after their 256-byte
     * aligned offset (pencil_update) */
    if (!HIPOK(hipHostMalloc((void **)&d->hcol, (size_t)(L.maxsub + L.ims + 8) * 4 * sizeof(zcx) + 512, 0))) {
        d->hcol = NULL;
        goto fail;
    }
    if (!HIPOK(hipHostMalloc((void **)&d->hgram, 2 * ss * sizeof(zcx), 0))) { d->hgram = NULL; goto fail; }
    if (!HIPOK(hipHostMalloc((void **)&d->hgco, 64, 0))) { d->hgco = NULL; goto fail; }
    if (!HIPOK(hipHostMalloc((void **)&d->hqzd, hqzd_bytes((size_t)L.maxsub), 0))) { d->hqzd = NULL; goto fail; }
    if (L.ims > 0 && o->callback_space == BBJD_CALLBACK_DEVICE) {
        d->look_slot = ((size_t)(L.ims + 1) * sizeof(zcx) + 255) / 256 * 256 + 256;
        if (!HIPOK(hipHostMalloc((void **)&d->hlook, GM_LOOK_MAX * d->look_slot, 0))) { d->hlook = NULL; goto fail; }
        for (; d->nlev < GM_LOOK_MAX; d->nlev++)
            if (!HIPOK(hipEventCreateWithFlags(&d->lev[d->nlev], hipEventDisableTiming |
                                                                      (o->blocking_sync ? hipEventBlockingSync : 0))))
                goto fail;
    }
    d->gx_cap = (int)L.ims + 1;
    if (!(d->gx = calloc((size_t)d->gx_cap, sizeof(hipGraphExec_t)))) goto fail;
    d->pair_res = calloc((size_t)L.nev + 1, sizeof(double));
    if (L.proj) {
        const size_t K = (size_t)L.nev + 1;
        /* H, H^-1 and the elimination's K x 2K work array */
        if (!(d->hh = calloc(4 * K * K, sizeof(zcx)))) goto fail;
        if (!HIPOK(hipHostMalloc((void **)&d->hpp, 3 * K * K * sizeof(c64), 0))) { d->hpp = NULL; goto fail; }
    }
    if (!ws->MA || !ws->MB || !ws->SA || !ws->SB || !ws->ZA || !ws->ZB || !ws->gmres_H || !ws->gmres_cs ||
        !ws->gmres_sn || !ws->gmres_g || !d->al || !d->be || !d->yv || !d->xv || !d->pair_res)
        goto fail;
    if (bbjd_qz_init(&d->qz, L.maxsub)) goto fail;
    if (o->blocking_sync &&
        !HIPOK(hipEventCreateWithFlags(&d->bev, hipEventDisableTiming | hipEventBlockingSync))) {
        d->bev = NULL;
        goto fail;
    }
    if (!HIPOK(hipStreamSynchronize(d->stream))) goto fail;
    return ws;
fail:
    workspace_free_impl(ws);
    return NULL;
}

bbjd_workspace *bbjd_workspace_alloc_gpu(const bbjd_params *p, const bbjd_gpu_options *o)
{
    pthread_rwlock_wrlock(&g_capture_rw);
    bbjd_workspace *ws = workspace_alloc_impl(p, o);
    if (ws && ws->dev->opt.graphs) __atomic_add_fetch(&g_graph_workspaces, 1, __ATOMIC_RELEASE);
    pthread_rwlock_unlock(&g_capture_rw);
    return ws;
}

bbjd_workspace *bbjd_workspace_alloc(const bbjd_params *p) { return bbjd_workspace_alloc_gpu(p, NULL); }

void bbjd_result_free(bbjd_result *r)
{
    if (!r) return;
    free(r->eigenvalues); free(r->eigenvectors); free(r->residuals);
    r->eigenvalues = NULL; r->eigenvectors = NULL; r->residuals = NULL;
}

void *bbjd_workspace_stream(const bbjd_workspace *ws) { return ws && ws->dev ? (void *)ws->dev->stream : NULL; }

int bbjd_workspace_stats(const bbjd_workspace *ws, bbjd_stats *st)
{
    if (!ws || !ws->dev || !st) return -1;
    *st = ws->dev->st;
    return 0;
}

void *bbjd_dev_alloc(size_t bytes)
{
    void *p = NULL;
    return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess ? p : NULL;
}

void bbjd_dev_free(void *p)
{
    if (p) hipFree(p);
}

int bbjd_dev_sync(void) { return HIPOK(hipDeviceSynchronize()) ? 0 : -1; }

int bbjd_time_op(bbjd_matvec_fn f, bbjd_int n, const void *x, void *y, void *data, int reps, double *avg_ms)
{
    hipEvent_t e0, e1;
    if (!f || reps <= 0 || !HIPOK(hipEventCreate(&e0))) return -1;
    if (!HIPOK(hipEventCreate(&e1))) { hipEventDestroy(e0); return -1; }
    int rc = 0;
    rc |= f(n, (const bbjd_zomplex *)x, (bbjd_zomplex *)y, data);       /* warm */
    rc |= !HIPOK(hipDeviceSynchronize());
    rc |= !HIPOK(hipEventRecord(e0, NULL));
    for (int r = 0; r < reps && !rc; r++) rc |= f(n, (const bbjd_zomplex *)x, (bbjd_zomplex *)y, data);
    rc |= !HIPOK(hipEventRecord(e1, NULL));
    float ms = 0;
    rc |= !HIPOK(hipEventSynchronize(e1));
    rc |= !HIPOK(hipEventElapsedTime(&ms, e0, e1));
    if (avg_ms) *avg_ms = ms / reps;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc ? -1 : 0;
}

int bbjd_memcpy_d2h(void *dst, const void *src, size_t bytes)
{ return HIPOK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)) ? 0 : -1; }
int bbjd_memcpy_h2d(void *dst, const void *src, size_t bytes)
{ return HIPOK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)) ? 0 : -1; }

/* ------------------------------------------------------------------ */
/* helpers                                                             */
/* ------------------------------------------------------------------ */
typedef struct {
    bbjd_workspace   *ws;
    struct bbjd_dev  *d;
    const bbjd_params *P;
    hipStream_t       s;
    bbjd_int          n, ld;
    bbjd_int          ms;      /* maxsub of this solve (restart threshold)   */
    bbjd_int          ldm;     /* leading dimension of the host k x k data   */
    zcx               nu0, mu0;  /* the test space's combination             */
    int               w_mixed;   /* a W column came from the Av fallback since
                                    the last restart: no triangular relation */
    int               graphs;    /* capture/replay the inner-solve segments  */
    int               pp;        /* this inner solve uses the projected
                                    preconditioner (BBJD_INNER_PROJECTED)   */
    char             *stage;     /* batched inner solve: the device slot that
                                    receives the Hessenberg columns and the
                                    scalar slots (NULL: copied to the host)  */
    size_t            stage_sc;  /* offset of the scalar slots in it         */
} ctx_t;

#define COLZ(M, j) ((void *)((zcx *)(M) + (size_t)(j) * c->ld))
#define COLC(M, j) ((void *)((c64 *)(M) + (size_t)(j) * c->ld))
#define COLZ_STATIC(M, j, ldv) ((void *)((zcx *)(M) + (size_t)(j) * (ldv)))

/* glibc rand() reproduced per solve: srand(s) == srandom(s) on the default
 * 128-byte TYPE_3 state; random_r on a private state of that size gives the
 * same sequence without touching the process-global generator. */
static void rng_seed(struct bbjd_dev *d, unsigned seed)
{
    memset(&d->rng, 0, sizeof(d->rng));
    initstate_r(seed, (char *)d->rng_state, sizeof(d->rng_state), &d->rng);
}
static double rng_unit(struct bbjd_dev *d)
{
    int32_t r;
    random_r(&d->rng, &r);
    return (double)r / 2147483647.0;   /* RAND_MAX */
}

static int sync_scal(ctx_t *c)
{
    struct bbjd_dev *d = c->d;
    if (!HIPOK(hipMemcpyAsync(d->hscal, d->rd.scal, SL_COUNT * 8, hipMemcpyDeviceToHost, c->s))) return -1;
    if (!HIPOK(dsync(c->d, c->s))) return -1;
    return 0;
}

/* Apply a user operator in the workspace's callback space. */
static int apply_op(ctx_t *c, bbjd_matvec_fn f, const void *x, void *y)
{
    struct bbjd_dev *d = c->d;
    if (d->opt.callback_space == BBJD_CALLBACK_DEVICE) {
        hipStream_t prev = tl_cb_stream;
        tl_cb_stream = c->s;
        int rc = f(c->n, (const bbjd_zomplex *)x, (bbjd_zomplex *)y, c->P->op_data);
        tl_cb_stream = prev;
        return rc;
    }
    const size_t b = (size_t)c->n * sizeof(zcx);
    if (!HIPOK(hipMemcpyAsync(d->hx, x, b, hipMemcpyDeviceToHost, c->s))) return -1;
    if (!HIPOK(dsync(c->d, c->s))) return -1;
    int rc = f(c->n, (const bbjd_zomplex *)d->hx, (bbjd_zomplex *)d->hy, c->P->op_data);
    if (rc) return rc;
    if (!HIPOK(hipMemcpyAsync(y, d->hy, b, hipMemcpyHostToDevice, c->s))) return -1;
    return 0;
}

/* ya = A x and yb = B x: one call of the fused callback when there is one
 * (device callbacks), else the two matvecs */
static int apply_ab(ctx_t *c, const void *x, void *ya, void *yb)
{
    struct bbjd_dev *d = c->d;
    if (d->opt.matvec_AB && d->opt.callback_space == BBJD_CALLBACK_DEVICE) {
        hipStream_t prev = tl_cb_stream;
        tl_cb_stream = c->s;
        int rc = d->opt.matvec_AB(c->n, (const bbjd_zomplex *)x, (bbjd_zomplex *)ya, (bbjd_zomplex *)yb,
                                  c->P->op_data);
        tl_cb_stream = prev;
        return rc;
    }
    int rc = apply_op(c, c->P->matvec_A, x, ya);
    return rc ? rc : apply_op(c, c->P->matvec_B, x, yb);
}

static int apply_pc(ctx_t *c, const void *x, void *y)
{
    if (c->P->precond) return apply_op(c, (bbjd_matvec_fn)c->P->precond, x, y);
    return HIPOK(hipMemcpyAsync(y, x, (size_t)c->n * sizeof(zcx), hipMemcpyDeviceToDevice, c->s)) ? 0 : -1;
}

static int upload_small(ctx_t *c, void *dst, const zcx *src, bbjd_int cnt)
{
    memcpy(c->d->hsmall, src, (size_t)cnt * sizeof(zcx));
    return HIPOK(hipMemcpyAsync(dst, c->d->hsmall, (size_t)cnt * sizeof(zcx), hipMemcpyHostToDevice, c->s)) ? 0
                                                                                                          : -1;
}

/* Column j of the deflation projectors (reference src/jd_core.c:516-528):
 * Q_right[:, j] = normalised CGS2(u), Q_left[:, j] = normalised
 * CGS2(nu0 Au + mu0 Bu); both normalisations only when the norm > 1e-300. */
static int deflation_append(ctx_t *c, bbjd_int j, const void *u, const void *Au, const void *Bu, zcx nu0,
                            zcx mu0)
{
    bbjd_workspace *ws = c->ws;
    struct bbjd_dev *d = c->d;
    int rc = 0;
    void *qr = COLZ(ws->Q_right, j), *ql = COLZ(ws->Q_left, j);
    rc |= !HIPOK(hipMemcpyAsync(qr, u, (size_t)c->n * 16, hipMemcpyDeviceToDevice, c->s));
    rc |= bbjd_k_cgs2_unit(c->s, c->n, (int)j, ws->Q_right, c->ld, qr, &d->rd, SL_Q);
    rc |= bbjd_k_lincomb(c->s, c->n, creal(nu0), cimag(nu0), Au, creal(mu0), cimag(mu0), Bu, ql);
    rc |= bbjd_k_cgs2_unit(c->s, c->n, (int)j, ws->Q_left, c->ld, ql, &d->rd, SL_Q);
    if (d->opt.precision == BBJD_PREC_MIXED) {
        rc |= bbjd_k_z2c(c->s, c->n, qr, COLC(d->qrc, j));
        rc |= bbjd_k_z2c(c->s, c->n, ql, COLC(d->qlc, j));
    }
    d->st.stream_bytes += 16.0 * c->n * (3 + 4.0 * j + 4 + 4.0 * j);
    return rc ? -1 : 0;
}

/* Givens with real cs >= 0 (reference src/jd_core.c:477-511). */
static void givens(zcx f, zcx g, double *cs, zcx *sn, zcx *r)
{
    double af = cabs(f), ag = cabs(g);
    if (ag < 1e-300) { *cs = 1.0; *sn = 0.0; *r = f; return; }
    if (af < 1e-300) { *cs = 0.0; *sn = conj(g) / ag; *r = ag; return; }
    double sc = af + ag, fs = af / sc, gs = ag / sc;
    double nrm = sc * sqrt(fs * fs + gs * gs);
    *cs = af / nrm;
    *sn = (f / af) * conj(g) / nrm;
    *r = (f / af) * nrm;
}

/* complex64 operator apply for the mixed inner solve: the c64 callback when
 * given, else the FP64 one through conversions (z64 scratch, n complex128). */
static int apply_c64(ctx_t *c, bbjd_op_c64_fn f64, bbjd_matvec_fn f, const void *x, void *y, void *z1, void *z2)
{
    if (f64 && c->d->opt.callback_space == BBJD_CALLBACK_DEVICE) {
        hipStream_t prev = tl_cb_stream;
        tl_cb_stream = c->s;
        int rc = f64(c->n, (const bbjd_cfloat *)x, (bbjd_cfloat *)y, c->P->op_data);
        tl_cb_stream = prev;
        return rc;
    }
    int rc = bbjd_k_c2z(c->s, c->n, x, z1);
    if (rc) return -1;
    if (f) rc = apply_op(c, f, z1, z2);
    else rc = !HIPOK(hipMemcpyAsync(z2, z1, (size_t)c->n * 16, hipMemcpyDeviceToDevice, c->s));
    if (rc) return rc;
    return bbjd_k_z2c(c->s, c->n, z2, y);
}

/* Captured segments of the inner solve (bbjd_gpu_options.graphs, device
 * callbacks).  seg_begin returns 1 when segment idx was replayed from its
 * graph (skip the body), 0 when the body is to run now (under capture when
 * graphs are on), -1 on error; seg_end closes a capture, instantiates and
 * launches it.  A body enqueues work only; its error ends the capture too.
 * A capture holds the process-wide capture lock shared, so that no other
 * thread allocates or frees device memory meanwhile; instantiating and
 * destroying graphs, which do, hold it exclusively. */
static int seg_begin(ctx_t *c, int idx)
{
    struct bbjd_dev *d = c->d;
    if (!c->graphs) return 0;
    if (d->gx[idx]) {
        if (!HIPOK(hipGraphLaunch(d->gx[idx], c->s))) return -1;
        d->st.graph_launches++;
        return 1;
    }
    bbjd_capture_lock();
    if (HIPOK(hipStreamBeginCapture(c->s, hipStreamCaptureModeRelaxed))) return 0;
    bbjd_capture_unlock();
    return -1;
}

static int seg_end(ctx_t *c, int idx, int rc)
{
    struct bbjd_dev *d = c->d;
    if (!c->graphs) return rc;
    hipGraph_t g = NULL;
    const hipError_t e = hipStreamEndCapture(c->s, &g);
    bbjd_capture_unlock();
    if (rc || e != hipSuccess) {
        if (e == hipSuccess && g) (void)hipGraphDestroy(g);
        return rc ? rc : -2;
    }
    /* instantiation allocates device memory: exclusive */
    hipGraphExec_t x = NULL;
    pthread_rwlock_wrlock(&g_capture_rw);
    const hipError_t ei = hipGraphInstantiateWithFlags(&x, g, 0);
    (void)hipGraphDestroy(g);
    pthread_rwlock_unlock(&g_capture_rw);
    if (ei != hipSuccess) return -2;
    d->gx[idx] = x;
    d->st.graph_captures++;
    if (!HIPOK(hipGraphLaunch(x, c->s))) return -2;
    d->st.graph_launches++;
    return 0;
}

/* Inverse of the K x K complex matrix A (column-major, leading dimension
 * lda) into X (column-major, leading dimension K) by Gauss-Jordan elimination
 * with partial pivoting; returns 1 when a pivot falls below 1e-10 of the
 * largest entry (numerically singular for this use), 0 otherwise. */
static int small_inverse(bbjd_int K, const zcx *A, bbjd_int lda, zcx *X, zcx *work)
{
    /* work: K x 2K, [A | I] row-reduced in place */
    const bbjd_int ldw = K;
    double amax = 0;
    for (bbjd_int j = 0; j < K; j++)
        for (bbjd_int i = 0; i < K; i++) {
            work[i + j * ldw] = A[i + j * lda];
            work[i + (K + j) * ldw] = i == j ? 1.0 : 0.0;
            if (cabs(A[i + j * lda]) > amax) amax = cabs(A[i + j * lda]);
        }
    if (!(amax > 0)) return 1;
    for (bbjd_int c = 0; c < K; c++) {
        bbjd_int p = c;
        for (bbjd_int i = c + 1; i < K; i++)
            if (cabs(work[i + c * ldw]) > cabs(work[p + c * ldw])) p = i;
        if (cabs(work[p + c * ldw]) < 1e-10 * amax) return 1;
        if (p != c)
            for (bbjd_int j = c; j < 2 * K; j++) {
                const zcx t = work[c + j * ldw];
                work[c + j * ldw] = work[p + j * ldw];
                work[p + j * ldw] = t;
            }
        const zcx ip = 1.0 / work[c + c * ldw];
        for (bbjd_int j = c; j < 2 * K; j++) work[c + j * ldw] *= ip;
        for (bbjd_int i = 0; i < K; i++) {
            if (i == c) continue;
            const zcx f = work[i + c * ldw];
            if (f == 0) continue;
            for (bbjd_int j = c; j < 2 * K; j++) work[i + j * ldw] -= f * work[c + j * ldw];
        }
    }
    for (bbjd_int j = 0; j < K; j++)
        for (bbjd_int i = 0; i < K; i++) X[i + j * K] = work[i + (K + j) * ldw];
    return 0;
}

/* Set-up of the projected-preconditioner inner solve (BBJD_INNER_PROJECTED)
 * for ndefl = K deflation columns.  Columns of Yh = M^-1 Q_left and the
 * rows and columns of H = Q_right^H Yh from pp_valid on are (re)computed in
 * complex64 (column K - 1 holds the current Ritz pair's projector columns,
 * which change every outer iteration; the earlier ones are locked pairs'),
 * H comes to the host, and H^-1 goes to the device.  One host sync.
 * Returns 0, 1 when H is numerically singular (the call takes the
 * reference form), -1 when a callback failed, -2 on an enqueue error. */
static int pp_prepare(ctx_t *c, bbjd_int K)
{
    struct bbjd_dev *d = c->d;
    const bbjd_params *P = c->P;
    const bbjd_int n = c->n, ldH = d->pp_ld;
    bbjd_int v0 = d->pp_valid < K ? d->pp_valid : K - 1;
    if (v0 < 0) v0 = 0;
    const bbjd_int nnew = K - v0;
    c64 *hcol = (c64 *)d->hnew, *hrow = (c64 *)d->hnew + (size_t)nnew * K;
    int rc = 0;
    for (bbjd_int j = v0; j < K; j++)
        if (apply_c64(c, d->opt.precond_c64, (bbjd_matvec_fn)P->precond, COLC(d->qlc, j), COLC(d->yc, j),
                      c->ws->tmp3, c->ws->tmp4))
            return -1;
    for (bbjd_int j = v0; j < K; j++) {
        /* H[:, j] = Q^H Yh[:, j] */
        rc |= bbjd_k_gemv_c_c(c->s, n, (int)K, d->qrc, c->ld, COLC(d->yc, j), NULL, hcol + (size_t)(j - v0) * K,
                              NULL, &d->rd);
        /* H[j, :v0] = conj(Yh[:, :v0]^H Q[:, j]) */
        if (v0 > 0)
            rc |= bbjd_k_gemv_c_c(c->s, n, (int)v0, d->yc, c->ld, COLC(d->qrc, j), NULL,
                                  hrow + (size_t)(j - v0) * v0, NULL, &d->rd);
    }
    const size_t cnt = (size_t)nnew * K + (size_t)nnew * v0;
    rc |= !HIPOK(hipMemcpyAsync(d->hpp, d->hnew, cnt * sizeof(c64), hipMemcpyDeviceToHost, c->s));
    if (rc || !HIPOK(dsync(c->d, c->s))) return -2;
    zcx *H = d->hh;
    const c64 *hc = d->hpp, *hr = d->hpp + (size_t)nnew * K;
    for (bbjd_int j = v0; j < K; j++) {
        for (bbjd_int i = 0; i < K; i++) {
            const c64 e = hc[(size_t)(j - v0) * K + i];
            H[i + j * ldH] = e.re + e.im * I;
        }
        for (bbjd_int i = 0; i < v0; i++) {
            const c64 e = hr[(size_t)(j - v0) * v0 + i];
            H[j + i * ldH] = e.re - e.im * I;
        }
    }
    d->pp_valid = K - 1;
    /* H^-1 (complex128 on the host, then complex64 to the device) */
    zcx *X = d->hh + (size_t)ldH * ldH;
    const int sing = small_inverse(K, H, ldH, X, X + (size_t)ldH * ldH);
    if (!sing) {
        c64 *up = d->hpp + (size_t)2 * K * K;
        for (size_t i = 0; i < (size_t)K * K; i++) { up[i].re = (float)creal(X[i]); up[i].im = (float)cimag(X[i]); }
        rc |= !HIPOK(hipMemcpyAsync(d->hinv, up, (size_t)K * K * sizeof(c64), hipMemcpyHostToDevice, c->s));
    }
    if (rc) return -2;
    return sing ? 1 : 0;
}

/* Copies of an Arnoldi step's results (and the prologue's scalars): to the
 * host (hh, hs), or in a batched inner solve to the staging slot, where the
 * column ends right before the scalar slots (at stage_sc).  hs non-NULL: the
 * column is not contiguous with the slots on the device and goes first. */
static int copy_column(ctx_t *c, void *hh, const void *hsum, size_t colb, const double *hs)
{
    if (!hs) return 0;
    if (c->stage) return bbjd_k_copy(c->s, c->stage + c->stage_sc - colb, hsum, colb) != 0;
    return !HIPOK(hipMemcpyAsync(hh, hsum, colb, hipMemcpyDeviceToHost, c->s));
}
static int copy_scalars(ctx_t *c, void *hh, const void *hsum, size_t colb, double *hs)
{
    struct bbjd_dev *d = c->d;
    if (c->stage) {
        if (hs) return bbjd_k_copy(c->s, c->stage + c->stage_sc, d->rd.scal, SL_COUNT * 8) != 0;
        return bbjd_k_copy(c->s, c->stage + c->stage_sc - colb, hsum, colb + SL_COUNT * 8) != 0;
    }
    if (hs) return !HIPOK(hipMemcpyAsync(hs, d->rd.scal, SL_COUNT * 8, hipMemcpyDeviceToHost, c->s));
    return !HIPOK(hipMemcpyAsync(hh, hsum, colb + SL_COUNT * 8, hipMemcpyDeviceToHost, c->s));
}

/* Enqueues Arnoldi step j of the inner solve: v_{j+1} = M^-1 (I - QL QL^H)
 * (theta_b A - theta_a B)(I - QR QR^H) v_j, its CGS2 against v_0 .. v_j, the
 * new vector normalised (also into the next step's input copy), and copies
 * of the Hessenberg column to hh and of the scalar slots to hs (host).
 * With hs NULL the column is written on the device right before the scalar
 * slots (the end of the pencil staging, free during the inner solve) and
 * one copy brings both to hh: the column, then the slots.  Only device
 * work: the step reads v_j, theta and the norms from device memory.
 * Returns 0, -1 when a callback failed, -2 on an enqueue error. */
static int arnoldi_enqueue(ctx_t *c, bbjd_int j, bbjd_int ndefl, const void *QR, const void *QL, void *V, void *t1,
                           void *t2, void *hh, double *hs)
{
    struct bbjd_dev *d = c->d;
    const bbjd_params *P = c->P;
    const bbjd_int n = c->n;
    const int mixed = d->opt.precision == BBJD_PREC_MIXED;
    const size_t esz = mixed ? 8 : 16;
    const void *qr = mixed ? d->qrc : QR, *ql = mixed ? d->qlc : QL;
    void *vj1 = (char *)V + (size_t)(j + 1) * c->ld * esz;
    const size_t colb = (size_t)(j + 1) * esz;
    void *hsum = hs ? d->rd.hsum : (void *)((char *)d->rd.scal - colb);
    int rc = 0, cb = 0;
    if (!mixed) {
        /* t1 = v_j: written with v_j's normalisation (the previous step's
         * last launch, or the prologue's) */
        rc |= bbjd_k_project_out(c->s, n, (int)ndefl, QR, c->ld, t1, &d->rd);
        if (!rc && apply_ab(c, t1, t2, vj1)) cb = 1;
        rc |= bbjd_k_lincomb_dev(c->s, n, d->gco, t2, vj1, t2);
        rc |= bbjd_k_project_out(c->s, n, (int)ndefl, QL, c->ld, t2, &d->rd);
        if (!rc && !cb && apply_pc(c, t2, vj1)) cb = 1;
        rc |= bbjd_k_cgs2(c->s, n, (int)(j + 1), V, c->ld, vj1, &d->rd, SL_G, hsum);
        rc |= copy_column(c, hh, hsum, colb, hs);
        rc |= bbjd_k_scale_inv2(c->s, n, vj1, t1, d->rd.scal, SL_G, 0);   /* t1: the next step's */
    } else if (c->pp) {
        /* projected preconditioner: v_j is in Q_right-perp already, so the
         * operator reads it in place; after M^-1 one oblique projection
         * (I - Yh H^-1 Q^H) replaces both sides' projections */
        const void *vj = (const char *)V + (size_t)j * c->ld * esz;
        void *bq = d->tc[1], *e = d->tc[2];
        if (d->opt.matvec_AB_c64 && d->opt.callback_space == BBJD_CALLBACK_DEVICE) {
            hipStream_t prev = tl_cb_stream;
            tl_cb_stream = c->s;
            if (d->opt.matvec_AB_c64(n, (const bbjd_cfloat *)vj, (bbjd_cfloat *)bq, (bbjd_cfloat *)e, P->op_data))
                cb = 1;
            tl_cb_stream = prev;
        } else {
            if (apply_c64(c, d->opt.matvec_A_c64, P->matvec_A, vj, bq, t1, t2)) cb = 1;
            if (!cb && apply_c64(c, d->opt.matvec_B_c64, P->matvec_B, vj, e, t1, t2)) cb = 1;
        }
        rc |= bbjd_k_lincomb_dev_c(c->s, n, d->gco, bq, e, bq);
        if (!rc && !cb && apply_c64(c, d->opt.precond_c64, (bbjd_matvec_fn)P->precond, bq, vj1, t1, t2)) cb = 1;
        rc |= bbjd_k_gemv_c_c(c->s, n, (int)ndefl, qr, c->ld, vj1, NULL, d->rd.coef, NULL, &d->rd);
        rc |= bbjd_k_update_pre_c(c->s, n, (int)ndefl, d->yc, c->ld, d->hinv, d->rd.coef, vj1);
        rc |= bbjd_k_cgs2_c(c->s, n, (int)(j + 1), V, c->ld, vj1, &d->rd, SL_G, hsum);
        rc |= copy_column(c, hh, hsum, colb, hs);
        rc |= bbjd_k_scale_inv_c(c->s, n, vj1, d->rd.scal, SL_G, 0);
    } else {
        void *a = d->tc[0], *bq = d->tc[1], *e = d->tc[2];
        /* a = v_j: written with v_j's normalisation */
        rc |= bbjd_k_project_out_c(c->s, n, (int)ndefl, qr, c->ld, a, &d->rd);
        if (d->opt.matvec_AB_c64 && d->opt.callback_space == BBJD_CALLBACK_DEVICE) {
            hipStream_t prev = tl_cb_stream;
            tl_cb_stream = c->s;
            if (!rc && d->opt.matvec_AB_c64(n, (const bbjd_cfloat *)a, (bbjd_cfloat *)bq, (bbjd_cfloat *)e,
                                            P->op_data))
                cb = 1;
            tl_cb_stream = prev;
        } else {
            if (!rc && apply_c64(c, d->opt.matvec_A_c64, P->matvec_A, a, bq, t1, t2)) cb = 1;
            if (!rc && !cb && apply_c64(c, d->opt.matvec_B_c64, P->matvec_B, a, e, t1, t2)) cb = 1;
        }
        rc |= bbjd_k_lincomb_dev_c(c->s, n, d->gco, bq, e, bq);
        rc |= bbjd_k_project_out_c(c->s, n, (int)ndefl, ql, c->ld, bq, &d->rd);
        if (P->precond) {
            if (!rc && !cb && apply_c64(c, d->opt.precond_c64, (bbjd_matvec_fn)P->precond, bq, vj1, t1, t2))
                cb = 1;
        } else
            rc |= bbjd_copy_async(vj1, bq, (size_t)n * 8, hipMemcpyDeviceToDevice, c->s) != 0;
        rc |= bbjd_k_cgs2_c(c->s, n, (int)(j + 1), V, c->ld, vj1, &d->rd, SL_G, hsum);
        rc |= copy_column(c, hh, hsum, colb, hs);
        rc |= bbjd_k_scale_inv2_c(c->s, n, vj1, a, d->rd.scal, SL_G, 0);   /* a: the next step's */
    }
    /* both norms down (the Hessenberg column came above, or comes with them) */
    rc |= copy_scalars(c, hh, hsum, colb, hs);
    return cb ? -1 : rc ? -2 : 0;
}

/* Host side of one inner GMRES call: Hessenberg matrix, Givens rotations
 * and the rotated right-hand side (workspace arrays). */
typedef struct {
    zcx    *H, *sn, *g;
    double *cs;
    bbjd_int ldh, ims;
    double  tol, beta0;
} gm_t;

static void gm_begin(ctx_t *c, gm_t *G, bbjd_int ims, double tol)
{
    bbjd_workspace *ws = c->ws;
    G->H = (zcx *)ws->gmres_H; G->sn = (zcx *)ws->gmres_sn; G->g = (zcx *)ws->gmres_g;
    G->cs = (double *)ws->gmres_cs;
    G->ldh = ws->inner_maxiter + 1;
    G->ims = ims;
    G->tol = tol;
    G->beta0 = 0;
    memset(G->H, 0, (size_t)G->ldh * ws->inner_maxiter * sizeof(zcx));
}

static void gm_start(gm_t *G, double beta0)
{
    memset(G->g, 0, (size_t)(G->ims + 1) * sizeof(zcx));
    G->g[0] = beta0;
    G->beta0 = beta0;
}

/* Device work of the GMRES prologue: v_0 = M^-1 (I - QL QL^H)(-r) (FP64 and
 * the complex64 reference form) or (I - Yh H^-1 Q^H) M^-1 (-r) (projected
 * form), normalised, ||v_0|| in scal[SL_G]; the scalar slots then go to the
 * host (hscal) or the staging slot.  *cb: a callback failed. */
static int gm_prologue(ctx_t *c, bbjd_int ndefl, const void *QL, const void *r, void *V, void *t1, void *t2, int *cb)
{
    struct bbjd_dev *d = c->d;
    const bbjd_params *P = c->P;
    const bbjd_int n = c->n;
    const int mixed = d->opt.precision == BBJD_PREC_MIXED;
    const size_t esz = mixed ? 8 : 16;
    void *v0 = V;
    const void *ql = mixed ? d->qlc : QL;
    int rc = 0;
    (void)esz;
    if (!mixed) {
        rc |= bbjd_k_lincomb(c->s, n, -1.0, 0.0, r, 0.0, 0.0, r, t1);
        rc |= bbjd_k_project_out(c->s, n, (int)ndefl, QL, c->ld, t1, &d->rd);
        if (!rc && apply_pc(c, t1, v0)) *cb = 1;
        rc |= bbjd_k_norm(c->s, n, v0, &d->rd, SL_G);
        /* also into t1 (free again), where step 0 projects it */
        rc |= bbjd_k_scale_inv2(c->s, n, v0, t1, d->rd.scal, SL_G, 0);
    } else if (c->pp) {
        /* v0 = (I - Yh H^-1 Q^H) M^-1 (-r), normalised */
        void *tc0 = d->tc[0];
        rc |= bbjd_k_z2c(c->s, n, r, tc0);
        rc |= bbjd_k_scale_c(c->s, n, tc0, -1.0, 0.0);
        if (!rc && apply_c64(c, d->opt.precond_c64, (bbjd_matvec_fn)P->precond, tc0, v0, t1, t2)) *cb = 1;
        rc |= bbjd_k_gemv_c_c(c->s, n, (int)ndefl, d->qrc, c->ld, v0, NULL, d->rd.coef, NULL, &d->rd);
        rc |= bbjd_k_update_pre_c(c->s, n, (int)ndefl, d->yc, c->ld, d->hinv, d->rd.coef, v0);
        rc |= bbjd_k_norm_c(c->s, n, v0, &d->rd, SL_G);
        rc |= bbjd_k_scale_inv_c(c->s, n, v0, d->rd.scal, SL_G, 0);
    } else {
        void *tc0 = d->tc[0];
        rc |= bbjd_k_z2c(c->s, n, r, tc0);
        rc |= bbjd_k_scale_c(c->s, n, tc0, -1.0, 0.0);
        rc |= bbjd_k_project_out_c(c->s, n, (int)ndefl, ql, c->ld, tc0, &d->rd);
        if (P->precond) {
            if (!rc && apply_c64(c, d->opt.precond_c64, (bbjd_matvec_fn)P->precond, tc0, v0, t1, t2)) *cb = 1;
        } else
            rc |= bbjd_copy_async(v0, tc0, (size_t)n * 8, hipMemcpyDeviceToDevice, c->s) != 0;
        rc |= bbjd_k_norm_c(c->s, n, v0, &d->rd, SL_G);
        /* also into tc0 (free again), where step 0 projects it */
        rc |= bbjd_k_scale_inv2_c(c->s, n, v0, tc0, d->rd.scal, SL_G, 0);
    }
    if (c->stage) rc |= bbjd_k_copy(c->s, c->stage + c->stage_sc, d->rd.scal, SL_COUNT * 8) != 0;
    else rc |= !HIPOK(hipMemcpyAsync(d->hscal, d->rd.scal, SL_COUNT * 8, hipMemcpyDeviceToHost, c->s));
    return rc;
}

/* Column j of the Arnoldi process arrived (hh: the Hessenberg column in the
 * basis precision, hs: the scalar slots): the reference's breakdown test,
 * Givens update and stopping rule (src/jd_core.c:674-706).  Returns 1 when
 * the solve stops after this column. */
static int gm_column(ctx_t *c, gm_t *G, bbjd_int j, bbjd_int ndefl, const void *hh, const double *hs)
{
    struct bbjd_dev *d = c->d;
    const int mixed = d->opt.precision == BBJD_PREC_MIXED;
    const size_t esz = mixed ? 8 : 16;
    const bbjd_int ldh = G->ldh;
    zcx *H = G->H, *sn = G->sn, *g = G->g;
    double *cs = G->cs;
    d->st.inner_steps++;
    d->st.stream_bytes += (double)esz * c->n * ((c->pp ? 3.0 : 8.0) * ndefl + 4.0 * (j + 1) + 8);
    if (!mixed)
        for (bbjd_int i = 0; i <= j; i++) H[i + j * ldh] = ((const zcx *)hh)[i];
    else {
        const c64 *hc = (const c64 *)hh;
        for (bbjd_int i = 0; i <= j; i++) H[i + j * ldh] = hc[i].re + hc[i].im * I;
    }
    const double wb = hs[SL_NB], wa = hs[SL_G];
    H[(j + 1) + j * ldh] = wa;
    if (wa < 1e-14 * wb) return 1;                           /* :674 */
    for (bbjd_int i = 0; i < j; i++) {                       /* :680-685 */
        zcx h1 = H[i + j * ldh], h2 = H[i + 1 + j * ldh];
        H[i + j * ldh] = cs[i] * h1 + sn[i] * h2;
        H[i + 1 + j * ldh] = -conj(sn[i]) * h1 + cs[i] * h2;
    }
    double cj; zcx sj, rj;
    givens(H[j + j * ldh], H[j + 1 + j * ldh], &cj, &sj, &rj);
    cs[j] = cj; sn[j] = sj;
    H[j + j * ldh] = rj;
    H[j + 1 + j * ldh] = 0.0;
    zcx g1 = g[j], g2 = g[j + 1];
    g[j] = cs[j] * g1 + sn[j] * g2;
    g[j + 1] = -conj(sn[j]) * g1 + cs[j] * g2;
    return cabs(g[j + 1]) < G->tol * G->beta0;
}

/* z = V y from the first m columns (back substitution on the host,
 * reference src/jd_core.c:709-716), projected with I - QR QR^H in FP64
 * (:724).  Only device work, through recordable launches and copies. */
static int gm_finish(ctx_t *c, gm_t *G, bbjd_int m, bbjd_int ndefl, const void *QR, const void *V, void *z)
{
    struct bbjd_dev *d = c->d;
    const bbjd_int n = c->n, ldh = G->ldh;
    const int mixed = d->opt.precision == BBJD_PREC_MIXED;
    zcx *H = G->H, *y = G->g;
    int rc = 0;
    for (bbjd_int i = m - 1; i >= 0; i--) {
        for (bbjd_int k2 = i + 1; k2 < m; k2++) y[i] -= H[i + k2 * ldh] * y[k2];
        if (cabs(H[i + i * ldh]) > 1e-300) y[i] /= H[i + i * ldh];
    }
    if (m == 0) {
        rc |= bbjd_memset_async(z, 0, (size_t)n * 16, c->s) != 0;
    } else if (!mixed) {
        memcpy(d->hsmall, y, (size_t)m * sizeof(zcx));
        rc |= bbjd_copy_async(d->rd.coef, d->hsmall, (size_t)m * sizeof(zcx), hipMemcpyHostToDevice, c->s) != 0;
        rc |= bbjd_k_gemv_n(c->s, n, (int)m, V, c->ld, d->rd.coef, z, &d->rd, SL_T1);
    } else {
        c64 *yc = (c64 *)d->hsmall;
        for (bbjd_int i = 0; i < m; i++) { yc[i].re = (float)creal(y[i]); yc[i].im = (float)cimag(y[i]); }
        rc |= bbjd_copy_async(d->rd.coef, yc, (size_t)m * 8, hipMemcpyHostToDevice, c->s) != 0;
        rc |= bbjd_k_gemv_n_c(c->s, n, (int)m, V, c->ld, d->rd.coef, d->tc[0], &d->rd, SL_T1);
        rc |= bbjd_k_c2z(c->s, n, d->tc[0], z);
    }
    rc |= bbjd_k_project_out(c->s, n, (int)ndefl, QR, c->ld, z, &d->rd);  /* :724, FP64 */
    return rc;
}

/* One inner solve submitted to a batcher (the workspace's thread waits). */
struct gm_job {
    ctx_t          *c;
    bbjd_int        ndefl, ims;
    const void     *QR, *QL, *r;
    void           *z;
    double          tol;
    double          t_queued;
    int             rc, done;
    struct gm_job  *next;
    /* batch state */
    gm_t            G;
    bbjd_int        m, enq;        /* steps consumed / enqueued          */
    int             active, cb;
    hipStream_t     own;
};

static int batcher_run(bbjd_batcher *B, struct gm_job *job);
static int batch_applies(const ctx_t *c);

/* Left-preconditioned GMRES on the projected correction equation
 * (reference src/jd_core.c:568-727).  r, z: device complex128; QR, QL: the
 * ndefl leading columns of the deflation projectors.  Arnoldi is CGS2 with
 * the same DGKS test and Hessenberg column (sum of both passes) as the
 * reference's MGS2; Givens, least squares and stopping rule are the
 * reference's, on the host in FP64.  Mixed precision: basis, operator
 * applies and projections in complex64 with FP32 accumulation.
 *
 * The device work of the prologue and of each Arnoldi step is one segment
 * that ends by copying the scalars (and the Hessenberg column) to the host:
 * theta_b and -theta_a are read from device memory (gco) and the new basis
 * vector is scaled by the reciprocal of its device-side norm, so with
 * graphs a segment is captured once per deflation count and replayed on
 * later outer iterations.  With a batcher the call is handed to it and runs
 * with other workspaces' calls (batcher_run). */
static int gmres(ctx_t *c, bbjd_int ndefl, const void *QR, const void *QL, zcx ta, zcx tb, const void *r, void *z,
                 bbjd_int ims, double tol, int allow_pp)
{
    bbjd_workspace *ws = c->ws;
    struct bbjd_dev *d = c->d;
    const bbjd_int n = c->n;
    const int mixed = d->opt.precision == BBJD_PREC_MIXED;
    int rc = 0, seg;
    /* FP64 temporaries: t1 = tmp3, t2 = tmp4 as in the reference (tmp4 holds
     * r, read only by the first step) */
    void *t1 = ws->tmp3, *t2 = ws->tmp4;
    void *V = mixed ? d->gvc : ws->gmres_V;
    const size_t esz = mixed ? 8 : 16;
#define GCOL(j) ((void *)((char *)V + (size_t)(j) * c->ld * esz))

    c->graphs = d->opt.graphs && d->opt.callback_space == BBJD_CALLBACK_DEVICE;
    c->pp = 0;
    if (allow_pp && mixed && d->yc && c->P->precond && ndefl <= d->pp_ld) {
        const int pr = pp_prepare(c, ndefl);
        if (pr == -1) return -1;
        if (pr < 0) return -2;
        c->pp = pr == 0;
        d->st.inner_projected++;
        if (pr == 1) d->st.inner_projected_fallbacks++;
    }
    if (c->graphs) {
        struct gkey key;
        memset(&key, 0, sizeof(key));
        key.ndefl = ndefl; key.n = n; key.ld = c->ld; key.r = r; key.z = z; key.qr = QR; key.ql = QL;
        key.op = c->P->op_data;
        key.pp = c->pp;
        if (memcmp(&key, &d->gkey, sizeof(key))) {
            rc |= !HIPOK(dsync(c->d, c->s));
            graphs_drop(d);
            d->gkey = key;
        }
    }
    d->hgco[0] = creal(tb); d->hgco[1] = cimag(tb); d->hgco[2] = -creal(ta); d->hgco[3] = -cimag(ta);
    rc |= !HIPOK(hipMemcpyAsync(d->gco, d->hgco, 4 * sizeof(double), hipMemcpyHostToDevice, c->s));
    if (rc) return -2;

    if (batch_applies(c)) {
        /* the batch runs on the batcher's stream: everything enqueued here
         * before (the projectors, r, theta) has finished first */
        if (!HIPOK(dsync(c->d, c->s))) return -2;
        struct gm_job job;
        memset(&job, 0, sizeof(job));
        job.c = c; job.ndefl = ndefl; job.ims = ims; job.QR = QR; job.QL = QL; job.r = r; job.z = z; job.tol = tol;
        return batcher_run((bbjd_batcher *)d->opt.inner_batcher, &job);
    }

    gm_t G;
    gm_begin(c, &G, ims, tol);
    int cb = 0;              /* a callback failed */
    if ((seg = seg_begin(c, 0)) < 0) return -2;
    if (seg == 0) {
        rc |= gm_prologue(c, ndefl, QL, r, GCOL(0), t1, t2, &cb);
        rc = seg_end(c, 0, rc | cb);
    }
    if (cb) return -1;
    if (rc || !HIPOK(dsync(c->d, c->s))) return -2;
    const double beta0 = d->hscal[SL_G];
    if (beta0 < 1e-300) {
        return HIPOK(hipMemsetAsync(z, 0, (size_t)n * 16, c->s)) ? 0 : -2;
    }
    gm_start(&G, beta0);

    /* Lookahead (device callbacks, no graphs): a step needs nothing from the
     * host, so up to `look` steps are enqueued before the first of them is
     * read back, and the host tests the reference's stopping rule on each
     * column as it arrives.  Steps enqueued past the stop are discarded:
     * they only wrote basis columns past m, the temporaries and reduction
     * scratch, so the result is the same as stepping one at a time. */
    const int look = !c->graphs && d->hlook ? g_gmres_look : 1;
    /* the column fits before the scalar slots: one copy per step */
    const int joint = (size_t)(ims + 1) * esz <= (size_t)((const char *)d->rd.scal - (const char *)d->pencil);
    bbjd_int j, e = 0, fail_at = -1;
    for (j = 0; j < ims; j++) {
        const void *hh = d->hsmall;
        const double *hs = d->hscal;
        if (look > 1) {
            for (; e < ims && e < j + look && fail_at < 0; e++) {
                char *slot = d->hlook + (size_t)(e % look) * d->look_slot;
                const int er = arnoldi_enqueue(c, e, ndefl, QR, QL, V, t1, t2, slot,
                                               joint ? NULL : (double *)(slot + d->look_slot - 256));
                if (er == -2) return -2;
                if (er == -1) { fail_at = e; break; }
                if (!HIPOK(hipEventRecord(d->lev[e % look], c->s))) return -2;
            }
            if (j == fail_at) return -1;
            if (!HIPOK(hipEventSynchronize(d->lev[j % look]))) return -2;
            hh = d->hlook + (size_t)(j % look) * d->look_slot;
            hs = (const double *)((const char *)hh + (joint ? (size_t)(j + 1) * esz : d->look_slot - 256));
        } else {
            if ((seg = seg_begin(c, (int)j + 1)) < 0) return -2;
            if (seg == 0) {
                const int er = arnoldi_enqueue(c, j, ndefl, QR, QL, V, t1, t2, d->hsmall, d->hscal);
                cb = er == -1;
                rc = seg_end(c, (int)j + 1, er ? 1 : 0);
                if (er == -2) rc = -2;
            }
            if (cb) return -1;
            if (rc || !HIPOK(dsync(c->d, c->s))) return -2;
        }
        if (gm_column(c, &G, j, ndefl, hh, hs)) { j++; break; }
    }
    const bbjd_int m = j;
    if (e > m) d->st.inner_steps_discarded += e - m;
#undef GCOL
    return gm_finish(c, &G, m, ndefl, QR, V, z) ? -2 : 0;
}

/* ------------------------------------------------------------------ */
/* inner-solve batcher                                                 */
/* ------------------------------------------------------------------ */
/* Threads with a HIP stream each run batches of inner solves handed over by
 * workspaces solving at once.  A batch records every solve's device work of
 * one phase (the prologue, Arnoldi step j, the final combination) through
 * the recording launchers, then launches all of them on the thread's stream,
 * kernels of the same instantiation from different solves merged into one
 * launch (bbjd_rec_flush); each solve's kernels run with its own grid,
 * arguments and order, so its results are those of running alone.  Each
 * step's Hessenberg columns and scalar slots come back in one copy of the
 * staging slots, and the host then runs each solve's Givens update and
 * stopping rule. */
#define BATCH_THREADS_MAX 8
#define BATCH_JOBS_MAX 16
#define BATCH_CAND_MAX 256
static int g_batch_pick = 1;
void bbjd_set_batch_pick(int nearest_n) { g_batch_pick = nearest_n != 0; }

struct bthread {
    bbjd_batcher *B;
    pthread_t     th;
    hipStream_t   s;
    hipEvent_t    ev;
    hipEvent_t    lev[GM_LOOK_MAX];   /* after each step in flight's copy */
    void         *rec;
    char         *dstage, *hstage;
    int           started;
};

struct bbjd_batcher {
    pthread_mutex_t    mu;
    pthread_cond_t     cv_job, cv_done;
    struct gm_job     *head, *tail;
    int                queued;     /* jobs waiting                         */
    int                running;    /* jobs in batches                      */
    int                active;     /* solves of the batcher's workspaces   */
    int                stop;
    int                nth, max_jobs;
    double             window;
    bbjd_int           ims_max;
    size_t             slot, stage_sc;
    const void        *busy[BATCH_THREADS_MAX * BATCH_JOBS_MAX];   /* operators in running batches */
    int                nbusy;
    struct bthread     t[BATCH_THREADS_MAX];
    bbjd_batcher_stats st;
};

static int batch_applies(const ctx_t *c)
{
    const struct bbjd_dev *d = c->d;
    const bbjd_batcher *B = (const bbjd_batcher *)d->opt.inner_batcher;
    return B && d->opt.precision == BBJD_PREC_MIXED && !c->pp && !c->graphs &&
           d->opt.callback_space == BBJD_CALLBACK_DEVICE && d->opt.matvec_AB_c64 == bbjd_op_matvec_AB_c64 &&
           d->opt.precond_c64 == bbjd_op_precond_c64 && c->P->op_data && c->ws->inner_maxiter <= B->ims_max;
}

static int batcher_run(bbjd_batcher *B, struct gm_job *job)
{
    pthread_mutex_lock(&B->mu);
    job->t_queued = now_s();
    job->next = NULL;
    if (B->tail) B->tail->next = job;
    else B->head = job;
    B->tail = job;
    B->queued++;
    pthread_cond_broadcast(&B->cv_job);
    while (!job->done) pthread_cond_wait(&B->cv_done, &B->mu);
    pthread_mutex_unlock(&B->mu);
    return job->rc;
}

static void batcher_enter(bbjd_batcher *B)
{
    pthread_mutex_lock(&B->mu);
    B->active++;
    pthread_mutex_unlock(&B->mu);
}

static void batcher_leave(bbjd_batcher *B)
{
    pthread_mutex_lock(&B->mu);
    B->active--;
    pthread_cond_broadcast(&B->cv_job);
    pthread_mutex_unlock(&B->mu);
}

/* launches the recorded phase, brings the staging slots of nj jobs to the
 * host and waits for both */
static int batch_phase_end(struct bthread *T, int nj, int copy)
{
    bbjd_batcher *B = T->B;
    long nl = 0;
    int rc = bbjd_rec_flush(T->rec, T->s, &nl);
    if (copy) rc |= !HIPOK(hipMemcpyAsync(T->hstage, T->dstage, (size_t)nj * B->slot, hipMemcpyDeviceToHost, T->s));
    rc |= !HIPOK(hipEventRecord(T->ev, T->s));
    rc |= !HIPOK(hipEventSynchronize(T->ev));
    return rc;
}

/* runs a batch; returns the Arnoldi steps consumed (counted under the
 * batcher's lock by the caller) */
static long batch_run(struct bthread *T, struct gm_job **jobs, int nj)
{
    bbjd_batcher *B = T->B;
    int err = 0;
    long steps = 0;
    for (int i = 0; i < nj; i++) {
        struct gm_job *J = jobs[i];
        ctx_t *c = J->c;
        J->own = c->s;
        c->s = T->s;
        c->stage = T->dstage + (size_t)i * B->slot;
        c->stage_sc = B->stage_sc;
        J->rc = 0;
        J->cb = 0;
        J->m = 0;
        J->enq = 0;
        J->active = 1;
        gm_begin(c, &J->G, J->ims, J->tol);
        bbjd_rec_begin(T->rec, i);
        const int r = gm_prologue(c, J->ndefl, J->QL, J->r, c->d->gvc, c->ws->tmp3, c->ws->tmp4, &J->cb);
        bbjd_rec_end();
        if (J->cb) { J->rc = -1; J->active = 0; }
        else if (r) { J->rc = -2; J->active = 0; }
    }
    err |= batch_phase_end(T, nj, 1);
    for (int i = 0; i < nj && !err; i++) {
        struct gm_job *J = jobs[i];
        if (!J->active) continue;
        const double beta0 = ((const double *)(T->hstage + (size_t)i * B->slot + B->stage_sc))[SL_G];
        if (beta0 < 1e-300) { J->active = 0; J->m = -1; continue; }   /* z = 0 */
        gm_start(&J->G, beta0);
    }
    /* Arnoldi steps with lookahead, as in the single solve: up to L steps
     * of every running solve are in flight, each step one flush (all
     * solves' kernels of that step) and one copy of the staging slots; a
     * solve's steps past its stop are discarded (they write only basis
     * columns past m, temporaries and reduction scratch) */
    const int L = g_gmres_look;
    bbjd_int e = 0;
    for (bbjd_int j = 0; !err; j++) {
        while (e < j + L && !err) {
            int na = 0;
            for (int i = 0; i < nj; i++) {
                struct gm_job *J = jobs[i];
                if (!J->active || e >= J->ims) continue;
                ctx_t *c = J->c;
                struct bbjd_dev *d = c->d;
                const int joint =
                    (size_t)(J->ims + 1) * 8 <= (size_t)((const char *)d->rd.scal - (const char *)d->pencil);
                c->stage = T->dstage + ((size_t)(e % L) * BATCH_JOBS_MAX + i) * B->slot;
                bbjd_rec_begin(T->rec, i);
                const int er = arnoldi_enqueue(c, e, J->ndefl, J->QR, J->QL, d->gvc, c->ws->tmp3, c->ws->tmp4,
                                               NULL, joint ? NULL : d->hscal);
                bbjd_rec_end();
                if (er) { J->rc = er == -1 ? -1 : -2; J->active = 0; continue; }
                J->enq = e + 1;
                na++;
            }
            if (!na) break;
            const size_t off = (size_t)(e % L) * BATCH_JOBS_MAX * B->slot;
            long nl = 0;
            err |= bbjd_rec_flush(T->rec, T->s, &nl);
            err |= !HIPOK(hipMemcpyAsync(T->hstage + off, T->dstage + off, (size_t)nj * B->slot,
                                         hipMemcpyDeviceToHost, T->s));
            err |= !HIPOK(hipEventRecord(T->lev[e % L], T->s));
            e++;
        }
        if (err || j >= e) break;
        err |= !HIPOK(hipEventSynchronize(T->lev[j % L]));
        if (err) break;
        for (int i = 0; i < nj; i++) {
            struct gm_job *J = jobs[i];
            if (!J->active || J->enq <= j) continue;
            const char *hb = T->hstage + ((size_t)(j % L) * BATCH_JOBS_MAX + i) * B->slot;
            const double *hs = (const double *)(hb + B->stage_sc);
            const void *hh = hb + B->stage_sc - (size_t)(j + 1) * 8;
            steps++;
            int stop = gm_column(J->c, &J->G, j, J->ndefl, hh, hs);
            if (stop || j + 1 == J->ims) {
                J->m = j + 1;
                J->active = 0;
                J->c->d->st.inner_steps_discarded += J->enq - J->m;
            }
        }
    }
    if (!err) {
        for (int i = 0; i < nj; i++) {
            struct gm_job *J = jobs[i];
            if (J->rc) continue;
            ctx_t *c = J->c;
            bbjd_rec_begin(T->rec, i);
            int r;
            if (J->m < 0) r = bbjd_memset_async(J->z, 0, (size_t)c->n * 16, c->s);
            else r = gm_finish(c, &J->G, J->m, J->ndefl, J->QR, c->d->gvc, J->z);
            bbjd_rec_end();
            if (r) J->rc = -2;
        }
        err |= batch_phase_end(T, nj, 0);
    }
    for (int i = 0; i < nj; i++) {
        struct gm_job *J = jobs[i];
        if (err && !J->rc) J->rc = -2;
        J->c->s = J->own;
        J->c->stage = NULL;
    }
    return steps;
}

static void *batcher_main(void *arg)
{
    struct bthread *T = (struct bthread *)arg;
    bbjd_batcher *B = T->B;
    struct gm_job *jobs[BATCH_JOBS_MAX];
    pthread_mutex_lock(&B->mu);
    for (;;) {
        while (!B->head && !B->stop) pthread_cond_wait(&B->cv_job, &B->mu);
        if (!B->head) break;
        /* more may come while solves of the batcher's workspaces are neither
         * waiting here nor in a batch: wait for them up to the window */
        while (B->head && B->queued < B->max_jobs && B->queued + B->running < B->active && !B->stop) {
            const double left = B->head->t_queued + B->window - now_s();
            if (left <= 0) break;
            struct timespec ts;
            clock_gettime(CLOCK_REALTIME, &ts);
            const double t = ts.tv_sec + ts.tv_nsec * 1e-9 + left;
            ts.tv_sec = (time_t)t;
            ts.tv_nsec = (long)((t - (double)ts.tv_sec) * 1e9);
            pthread_cond_timedwait(&B->cv_job, &B->mu, &ts);
        }
        if (!B->head) continue;
        /* take up to max_jobs, one per operator and none whose operator is in
         * a running batch (an operator's scratch serves one apply at a time):
         * the longest-waiting one, then (g_batch_pick 1) those nearest to it
         * in n, whose launch sequences match best (the preconditioner's level
         * structure follows n), or (0) the next ones in arrival order */
        int nj = 0, ncand = 0;
        struct gm_job *cand[BATCH_CAND_MAX];
        for (struct gm_job *J = B->head; J && ncand < BATCH_CAND_MAX; J = J->next) {
            const void *op = J->c->P->op_data;
            int dup = 0;
            for (int i = 0; i < B->nbusy; i++) dup |= B->busy[i] == op;
            for (int i = 0; i < ncand; i++) dup |= cand[i]->c->P->op_data == op;
            if (!dup) cand[ncand++] = J;
        }
        if (g_batch_pick && ncand > B->max_jobs) {
            /* insertion sort of cand[1..] by distance in log n from cand[0] */
            const double l0 = log((double)cand[0]->c->n);
            for (int i = 2; i < ncand; i++) {
                struct gm_job *x = cand[i];
                const double dx = fabs(log((double)x->c->n) - l0);
                int k = i - 1;
                while (k >= 1 && fabs(log((double)cand[k]->c->n) - l0) > dx) { cand[k + 1] = cand[k]; k--; }
                cand[k + 1] = x;
            }
        }
        for (int i = 0; i < ncand && nj < B->max_jobs; i++) {
            cand[i]->done = 2;        /* taken (done is 0 while queued) */
            B->busy[B->nbusy++] = cand[i]->c->P->op_data;
            jobs[nj++] = cand[i];
        }
        {
            struct gm_job *prev = NULL;
            for (struct gm_job *J = B->head; J;) {
                struct gm_job *next = J->next;
                if (J->done == 2) {
                    J->done = 0;
                    if (prev) prev->next = next;
                    else B->head = next;
                    if (B->tail == J) B->tail = prev;
                } else
                    prev = J;
                J = next;
            }
        }
        if (!nj) {
            /* every waiting solve's operator is busy: wait for a batch to end */
            pthread_cond_wait(&B->cv_job, &B->mu);
            continue;
        }
        B->queued -= nj;
        B->running += nj;
        pthread_mutex_unlock(&B->mu);
        const long steps = batch_run(T, jobs, nj);
        pthread_mutex_lock(&B->mu);
        B->st.steps += steps;
        for (int i = 0; i < nj; i++) {
            jobs[i]->done = 1;
            for (int b = 0; b < B->nbusy; b++)
                if (B->busy[b] == jobs[i]->c->P->op_data) { B->busy[b] = B->busy[--B->nbusy]; break; }
        }
        B->running -= nj;
        B->st.batches++;
        B->st.jobs += nj;
        pthread_cond_broadcast(&B->cv_done);
        pthread_cond_broadcast(&B->cv_job);
    }
    pthread_mutex_unlock(&B->mu);
    return NULL;
}

bbjd_batcher *bbjd_batcher_create(int threads, int max_jobs, double window_us, bbjd_int inner_maxiter)
{
    if (threads < 1 || threads > BATCH_THREADS_MAX || max_jobs < 1 || max_jobs > BATCH_JOBS_MAX ||
        inner_maxiter < 1 || window_us < 0)
        return NULL;
    bbjd_batcher *B = (bbjd_batcher *)calloc(1, sizeof(*B));
    if (!B) return NULL;
    pthread_mutex_init(&B->mu, NULL);
    pthread_cond_init(&B->cv_job, NULL);
    pthread_cond_init(&B->cv_done, NULL);
    B->nth = threads;
    B->max_jobs = max_jobs;
    B->window = window_us * 1e-6;
    B->ims_max = inner_maxiter;
    /* a slot: the Hessenberg column (complex64, ending at stage_sc), then
     * the scalar slots */
    B->stage_sc = ((size_t)(inner_maxiter + 1) * 8 + 255) / 256 * 256;
    B->slot = B->stage_sc + 256;
    int ok = 1;
    for (int i = 0; i < threads && ok; i++) {
        struct bthread *T = &B->t[i];
        T->B = B;
        ok &= HIPOK(hipStreamCreateWithFlags(&T->s, hipStreamNonBlocking));
        ok &= HIPOK(hipEventCreateWithFlags(&T->ev, hipEventDisableTiming | hipEventBlockingSync));
        for (int l = 0; l < GM_LOOK_MAX; l++)
            ok &= HIPOK(hipEventCreateWithFlags(&T->lev[l], hipEventDisableTiming | hipEventBlockingSync));
        ok &= HIPOK(hipMalloc((void **)&T->dstage, (size_t)GM_LOOK_MAX * BATCH_JOBS_MAX * B->slot));
        ok &= HIPOK(hipHostMalloc((void **)&T->hstage, (size_t)GM_LOOK_MAX * BATCH_JOBS_MAX * B->slot,
                                  hipHostMallocDefault));
        T->rec = bbjd_rec_create();
        ok &= T->rec != NULL;
        if (ok) T->started = pthread_create(&T->th, NULL, batcher_main, T) == 0;
        ok &= T->started;
    }
    if (!ok) { bbjd_batcher_destroy(B); return NULL; }
    return B;
}

void bbjd_batcher_destroy(bbjd_batcher *B)
{
    if (!B) return;
    pthread_mutex_lock(&B->mu);
    B->stop = 1;
    pthread_cond_broadcast(&B->cv_job);
    pthread_mutex_unlock(&B->mu);
    for (int i = 0; i < B->nth; i++) {
        struct bthread *T = &B->t[i];
        if (T->started) pthread_join(T->th, NULL);
        if (T->rec) bbjd_rec_destroy(T->rec);
        if (T->hstage) (void)hipHostFree(T->hstage);
        if (T->dstage) (void)hipFree(T->dstage);
        if (T->ev) (void)hipEventDestroy(T->ev);
        for (int l = 0; l < GM_LOOK_MAX; l++)
            if (T->lev[l]) (void)hipEventDestroy(T->lev[l]);
        if (T->s) (void)hipStreamDestroy(T->s);
    }
    pthread_cond_destroy(&B->cv_job);
    pthread_cond_destroy(&B->cv_done);
    pthread_mutex_destroy(&B->mu);
    free(B);
}

void bbjd_batcher_get_stats(const bbjd_batcher *B, bbjd_batcher_stats *st)
{
    if (!B || !st) return;
    pthread_mutex_lock((pthread_mutex_t *)&B->mu);
    *st = B->st;
    st->recorded = st->launches = 0;
    for (int i = 0; i < B->nth; i++) {
        long r = 0, l = 0;
        if (B->t[i].rec) bbjd_rec_counts(B->t[i].rec, &r, &l);
        st->recorded += r;
        st->launches += l;
    }
    pthread_mutex_unlock((pthread_mutex_t *)&B->mu);
}

/* Pencil column/row k after expansion (reference src/jd_core.c:1085-1104):
 * MA[0:k+1, k] = W^H AV[:, k], MB likewise, MA[k, 0:k] = W[:, k]^H AV[:, :k].
 * Rows come as conj(AV^H w).  One host sync. */
/* New column and row of the projected pencil (reference src/jd_core.c:1085-1104):
 * MA[:k+1, k] = W^H AV[:, k], MB[:k+1, k] = W^H BV[:, k] in one two-vector pass;
 * the row k from one pass.  W[:, k] is the normalised CGS2 remainder of
 * nu0 AV[:, k] + mu0 BV[:, k] against W[:, :k], so C = nu0 MA + mu0 MB is
 * upper triangular (w_k is orthogonal to nu0 AV[:, j] + mu0 BV[:, j] =
 * W[:, :j+1] c_j for j < k, to the orthogonality of W): with nu0 != 0 the
 * row of MA is -(mu0 / nu0) times the row of MB, which alone is projected
 * (BV^H w_k); with nu0 = 0, C = MB and its row is zero, and the row of MA
 * is projected.  rowrel = 0 (some column of W since the last restart came
 * from the Av fallback, so the relation does not hold) projects both rows as
 * the reference does. */
static int pencil_update(ctx_t *c, bbjd_int k, int rowrel, int col_done)
{
    bbjd_workspace *ws = c->ws;
    struct bbjd_dev *d = c->d;
    const bbjd_int ms = c->ldm;
    zcx *pen = (zcx *)d->pencil;
    const int via_b = rowrel && cabs(c->nu0) > 0;
    const int via_a = rowrel && !via_b;
    int rc = 0;
    if (!col_done)
        rc |= bbjd_k_gemv_c(c->s, c->n, (int)(k + 1), ws->W, c->ld, COLZ(ws->AV, k), COLZ(ws->BV, k), pen,
                            pen + (ms + 1), &d->rd);
    if (k > 0) {
        if (!via_b)
            rc |= bbjd_k_gemv_c(c->s, c->n, (int)k, ws->AV, c->ld, COLZ(ws->W, k), NULL, pen + 2 * (ms + 1), NULL,
                                &d->rd);
        if (!via_a)
            rc |= bbjd_k_gemv_c(c->s, c->n, (int)k, ws->BV, c->ld, COLZ(ws->W, k), NULL, pen + 3 * (ms + 1), NULL,
                                &d->rd);
    }
    /* one copy: the pencil column/row buffer and, right after it, the scalar
     * slots (the norms of the expansion, for the breakdown checks) */
    const size_t gap = (size_t)((const char *)d->rd.scal - (const char *)pen);
    rc |= !HIPOK(hipMemcpyAsync(d->hcol, pen, gap + SL_COUNT * 8, hipMemcpyDeviceToHost, c->s));
    rc |= !HIPOK(dsync(c->d, c->s));
    if (rc) return -1;
    memcpy(d->hscal, (const char *)d->hcol + gap, SL_COUNT * 8);
    zcx *MA = (zcx *)ws->MA, *MB = (zcx *)ws->MB;
    const zcx *ca = d->hcol, *cb = d->hcol + (ms + 1), *ra = d->hcol + 2 * (ms + 1), *rb = d->hcol + 3 * (ms + 1);
    for (bbjd_int i = 0; i <= k; i++) { MA[i + k * ms] = ca[i]; MB[i + k * ms] = cb[i]; }
    if (via_b) {
        const zcx f = -c->mu0 / c->nu0;
        for (bbjd_int i = 0; i < k; i++) { MB[k + i * ms] = conj(rb[i]); MA[k + i * ms] = f * MB[k + i * ms]; }
    } else if (via_a) {
        for (bbjd_int i = 0; i < k; i++) { MA[k + i * ms] = conj(ra[i]); MB[k + i * ms] = 0.0; }
    } else {
        for (bbjd_int i = 0; i < k; i++) { MA[k + i * ms] = conj(ra[i]); MB[k + i * ms] = conj(rb[i]); }
    }
    d->st.stream_bytes += 16.0 * c->n * ((col_done ? 3.0 : (k + 3.0)) + (rowrel ? 1.0 : 2.0) * (k + 1));
    return 0;
}

/* One Cholesky-QR pass on W[:, :keep] in place: G = W^H W on the device,
 * G = R^H R on the host, W <- W R^-1 on the device.  Returns 1 (W unchanged)
 * when the pass cannot be trusted -- G not numerically positive definite, R
 * too ill-conditioned (diagonal ratio below 1e-6, where the loss of
 * orthogonality eps cond^2 would no longer be repaired by a second pass), or,
 * with check, W already far from orthonormal (max |G - I| > 0.1) -- so the
 * caller falls back to column-by-column CGS2. */
static int cholqr_pass(ctx_t *c, bbjd_int keep, int check)
{
    bbjd_workspace *ws = c->ws;
    struct bbjd_dev *d = c->d;
    const bbjd_int n = c->n;
    zcx *G = d->hgram;
    int rc = bbjd_k_gram(c->s, n, (int)keep, ws->W, c->ld, (int)keep, ws->W, c->ld, d->gram, keep, &d->rd);
    rc |= !HIPOK(hipMemcpyAsync(G, d->gram, (size_t)keep * keep * 16, hipMemcpyDeviceToHost, c->s));
    rc |= !HIPOK(dsync(c->d, c->s));
    if (rc) return -1;
    if (check)
        for (bbjd_int j = 0; j < keep; j++)
            for (bbjd_int i = 0; i <= j; i++)
                if (cabs(G[i + j * keep] - (i == j ? 1.0 : 0.0)) > 0.1) return 1;
    blas_int bk = (blas_int)keep, info = 0;
    scipy_zpotrf_("U", &bk, G, &bk, &info);
    if (info != 0) return 1;
    double dmin = INFINITY, dmax = 0;
    for (bbjd_int j = 0; j < keep; j++) {
        const double r = creal(G[j + j * keep]);
        if (r < dmin) dmin = r;
        if (r > dmax) dmax = r;
    }
    if (!(dmin > 1e-6 * dmax)) return 1;
    scipy_ztrtri_("U", "N", &bk, G, &bk, &info);
    if (info != 0) return 1;
    /* row-major R^-1 for the rotation: Zr[j keep + c] = (R^-1)[j, c] */
    for (bbjd_int j = 0; j < keep; j++)
        for (bbjd_int cc = 0; cc < keep; cc++) d->zrh[(size_t)j * keep + cc] = j <= cc ? G[j + cc * keep] : 0.0;
    rc |= !HIPOK(hipMemcpyAsync(d->zr, d->zrh, (size_t)keep * keep * 16, hipMemcpyHostToDevice, c->s));
    rc |= bbjd_k_rotate(c->s, n, (int)keep, (int)keep, ws->W, c->ld, d->zr);
    return rc ? -1 : 0;
}

/* Thick restart (reference src/jd_core.c:975-1030): V, AV, BV <- (.) ZB[:, :keep]
 * in place on the device; W[:, :keep] = orth(nu0 AV + mu0 BV); MA, MB =
 * W^H AV, W^H BV.
 *
 * The reference orthonormalises the new W column by column (CGS2 against the
 * previous columns, scaled when the norm exceeds 1e-300): the Q factor, with
 * positive real R diagonal, of X = nu0 AV + mu0 BV.  Here the same Q factor
 * comes from Cholesky QR twice -- two Gram passes and two in-place
 * triangular rotations instead of keep^2 / 2 vector passes -- and the
 * column-by-column CGS2 is kept as the fallback whenever a pass cannot be
 * trusted (near-dependent columns).  The pencil is two Gram passes. */
static int thick_restart(ctx_t *c, bbjd_int k, bbjd_int keep, zcx nu0, zcx mu0)
{
    bbjd_workspace *ws = c->ws;
    struct bbjd_dev *d = c->d;
    const bbjd_int ms = c->ldm, n = c->n;
    const zcx *ZB = (const zcx *)ws->ZB;
    int rc = 0;
    if (!d->zr_ready) {        /* else the device step wrote it (Schur vectors of C^-1 MB) */
        for (bbjd_int j = 0; j < k; j++)
            for (bbjd_int cc = 0; cc < keep; cc++) d->zrh[(size_t)j * keep + cc] = ZB[j + cc * ms];
        rc |= !HIPOK(hipMemcpyAsync(d->zr, d->zrh, (size_t)k * keep * 16, hipMemcpyHostToDevice, c->s));
    }
    d->zr_ready = 0;
    rc |= bbjd_k_rotate(c->s, n, (int)k, (int)keep, ws->V, c->ld, d->zr);
    rc |= bbjd_k_rotate(c->s, n, (int)k, (int)keep, ws->AV, c->ld, d->zr);
    rc |= bbjd_k_rotate(c->s, n, (int)k, (int)keep, ws->BV, c->ld, d->zr);
    /* X = nu0 AV + mu0 BV over the keep columns at once (padding rows
     * between columns are never read) */
    rc |= bbjd_k_lincomb(c->s, (keep - 1) * c->ld + n, creal(nu0), cimag(nu0), ws->AV, creal(mu0), cimag(mu0),
                         ws->BV, ws->W);
    if (rc) return -1;
    int q = cholqr_pass(c, keep, 0);
    if (q == 0) q = cholqr_pass(c, keep, 1);
    if (q < 0) return -1;
    if (q > 0) {
        /* column-by-column CGS2 on the current W (X, or X R1^-1 after a first
         * pass: the same column space, hence the same Q factor) */
        d->st.restart_fallbacks++;
        for (bbjd_int j = 0; j < keep; j++) {
            void *wj = COLZ(ws->W, j);
            rc |= bbjd_k_cgs2_unit(c->s, n, (int)j, ws->W, c->ld, wj, &d->rd, SL_T0);
        }
    }
    zcx *G = (zcx *)d->gram, *hg = d->hgram;
    rc |= bbjd_k_gram(c->s, n, (int)keep, ws->W, c->ld, (int)keep, ws->AV, c->ld, G, keep, &d->rd);
    rc |= bbjd_k_gram(c->s, n, (int)keep, ws->W, c->ld, (int)keep, ws->BV, c->ld, G + keep * keep, keep, &d->rd);
    rc |= !HIPOK(hipMemcpyAsync(hg, G, (size_t)2 * keep * keep * 16, hipMemcpyDeviceToHost, c->s));
    rc |= !HIPOK(dsync(c->d, c->s));
    zcx *MA = (zcx *)ws->MA, *MB = (zcx *)ws->MB;
    for (bbjd_int j = 0; j < keep; j++)
        for (bbjd_int i = 0; i < keep; i++) {
            MA[i + j * ms] = hg[i + j * keep];
            MB[i + j * ms] = hg[keep * keep + i + j * keep];
        }
    c->w_mixed = 0;     /* W is the orthonormal factor of nu0 AV + mu0 BV again */
    d->st.restarts++;
    d->st.stream_bytes += 16.0 * n * (3.0 * (k + keep) + 3.0 * keep + 4.0 * keep + 3.0 * keep);
    return rc ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* the outer iteration                                                 */
/* ------------------------------------------------------------------ */
/* Random start vector for a breakdown (reference src/jd_core.c:1043-1052):
 * entries rand()/RAND_MAX - 0.5 in both parts, uploaded to v. */
static int random_vector(ctx_t *c, void *v)
{
    struct bbjd_dev *d = c->d;
    const bbjd_int n = c->n;
    zcx *tmpv = d->hx ? d->hx : malloc((size_t)n * sizeof(zcx));
    if (!tmpv) return -2;
    for (bbjd_int i = 0; i < n; i++) {
        double re = rng_unit(d) - 0.5;
        double im = rng_unit(d) - 0.5;
        tmpv[i] = re + im * I;
    }
    int rc = !HIPOK(hipMemcpyAsync(v, tmpv, (size_t)n * 16, hipMemcpyHostToDevice, c->s));
    if (!d->hx) {
        rc |= !HIPOK(dsync(c->d, c->s));
        free(tmpv);
    }
    return rc ? -2 : 0;
}

/* Expansion of the search and test spaces by column k (reference
 * src/jd_core.c:1032-1104): v = CGS2(V, z) / ||.||, with a random vector
 * when ||v|| < 1e-14; AV, BV columns; w = CGS2(W, nu0 Av + mu0 Bv) / ||.||,
 * with Av in place of the combination when ||w|| < 1e-14; then the new row
 * and column of the projected pencil.
 *
 * With device-space callbacks the breakdown tests are deferred: the
 * expansion runs through with the norms applied on the device, and the norms
 * come back to the host with the pencil column (one synchronisation instead
 * of three).  In the rare breakdown case the column is recomputed along the
 * reference's branch; every result of the first attempt lives in column k
 * and is overwritten.  Returns 0, -1 callback error, -2 device error, -3
 * breakdown. */
static int expand(ctx_t *c, bbjd_int k, const void *z, zcx nu0, zcx mu0)
{
    bbjd_workspace *ws = c->ws;
    struct bbjd_dev *d = c->d;
    const bbjd_int n = c->n;
    void *vn = COLZ(ws->V, k), *wn = COLZ(ws->W, k), *an = COLZ(ws->AV, k), *bn = COLZ(ws->BV, k);
    const int optimistic = d->opt.callback_space == BBJD_CALLBACK_DEVICE;
    int rc = 0;
    /* z may already be column k (formed in place); a breakdown of v replaces
     * it with a random vector, so z is not needed after CGS2 */
    const int in_place = z == vn;
    if (!in_place) rc |= !HIPOK(hipMemcpyAsync(vn, z, (size_t)n * 16, hipMemcpyDeviceToDevice, c->s));
    double nrm;
    if (optimistic) {
        rc |= bbjd_k_cgs2_unit(c->s, n, (int)k, ws->V, c->ld, vn, &d->rd, SL_V);
        if (rc) return -2;
        if (apply_ab(c, vn, an, bn)) return -1;
        /* fused: w = nu0 Av + mu0 Bv is formed inside the first passes over W,
         * which also give the pencil column */
        zcx *pen = (zcx *)d->pencil;
        const int pc = bbjd_k_cgs2_unit_pencil(c->s, n, (int)k, ws->W, c->ld, wn, an, bn, creal(nu0), cimag(nu0),
                                               creal(mu0), cimag(mu0), pen, pen + (c->ldm + 1), &d->rd, SL_W);
        if (pc < 0) return -2;
        if (pc > 0) {
            rc |= bbjd_k_lincomb(c->s, n, creal(nu0), cimag(nu0), an, creal(mu0), cimag(mu0), bn, wn);
            rc |= bbjd_k_cgs2_unit(c->s, n, (int)k, ws->W, c->ld, wn, &d->rd, SL_W);
        }
        if (rc || pencil_update(c, k, !c->w_mixed, pc == 0)) return -2;
        d->st.stream_bytes += 16.0 * n * (2.0 * k + 4 + 2.0 * k + 2 + (pc == 0 ? 2 : 3));
        const double nv = d->hscal[SL_V], nw = d->hscal[SL_W];
        if (nv >= 1e-14 && nw >= 1e-14) return 0;
        if (nv >= 1e-14) {
            /* only w broke down: v, Av, Bv stand; w from Av (the reference's
             * fallback), then the pencil again */
            rc |= !HIPOK(hipMemcpyAsync(wn, an, (size_t)n * 16, hipMemcpyDeviceToDevice, c->s));
            rc |= bbjd_k_cgs2_unit(c->s, n, (int)k, ws->W, c->ld, wn, &d->rd, SL_W);
            c->w_mixed = 1;
            if (rc || pencil_update(c, k, 0, 0)) return -2;
            return 0;
        }
        nrm = nv;           /* v broke down: the random-vector branch below */
    } else {
        rc |= bbjd_k_cgs2(c->s, n, (int)k, ws->V, c->ld, vn, &d->rd, SL_V, NULL);
        if (rc || sync_scal(c)) return -2;
        nrm = d->hscal[SL_V];
    }
    if (nrm < 1e-14) {
        if (random_vector(c, vn)) return -2;
        rc |= bbjd_k_cgs2(c->s, n, (int)k, ws->V, c->ld, vn, &d->rd, SL_V, NULL);
        if (rc || sync_scal(c)) return -2;
        nrm = d->hscal[SL_V];
        if (nrm < 1e-14) return -3;
    }
    rc |= bbjd_k_scale(c->s, n, vn, 1.0 / nrm, 0.0);
    if (rc) return -2;
    if (apply_ab(c, vn, an, bn)) return -1;
    rc |= bbjd_k_lincomb(c->s, n, creal(nu0), cimag(nu0), an, creal(mu0), cimag(mu0), bn, wn);
    rc |= bbjd_k_cgs2(c->s, n, (int)k, ws->W, c->ld, wn, &d->rd, SL_W, NULL);
    if (rc || sync_scal(c)) return -2;
    const int w_fallback = d->hscal[SL_W] < 1e-14;
    if (w_fallback) {
        rc |= !HIPOK(hipMemcpyAsync(wn, an, (size_t)n * 16, hipMemcpyDeviceToDevice, c->s));
        rc |= bbjd_k_cgs2(c->s, n, (int)k, ws->W, c->ld, wn, &d->rd, SL_W, NULL);
    }
    rc |= bbjd_k_scale_inv(c->s, n, wn, d->rd.scal, SL_W, 1);
    d->st.stream_bytes += 16.0 * n * (2.0 * k + 4 + 3 + 2.0 * k + 2);
    if (w_fallback) c->w_mixed = 1;
    if (rc || pencil_update(c, k, !c->w_mixed, 0)) return -2;
    return 0;
}

/* The projected step on the device (qz_mode BBJD_QZ_DEVICE): the ordered
 * Schur form of C^-1 MB, the reference's selection and target, the Ritz
 * coefficients in d->xv and, with keep > 0 (a restart follows), the leading
 * keep Schur vectors as the restart's rotation in d->zr (bbjd_qz_device_batch,
 * one problem on the solve's stream).  Returns 0 with *p, *ta, *tb set; 1
 * when the host step is to decide this iteration (k > 64, nu0 = 0, a w from
 * the Av fallback since the last restart, or a status other than OK from the
 * device: C not triangular, accuracy, a tie); -1 on a device error. */
static int qz_device(ctx_t *c, bbjd_int k, const zcx *MA, const zcx *MB, const zcx *conv, bbjd_int nconv,
                     bbjd_int nkeep, bbjd_int keep, bbjd_int *p, zcx *ta, zcx *tb)
{
    struct bbjd_dev *d = c->d;
    const bbjd_int ldm = c->ldm, msw = c->ws->maxsub, nevw = c->ws->nev;
    if (k > 64 || cabs(c->nu0) == 0 || c->w_mixed) return 1;
    char *db = (char *)d->qzd;
    const size_t mb = (size_t)msw * msw * 16, cb = (size_t)k * ldm * 16;
    const size_t o_conv = 2 * mb, o_al = o_conv + (size_t)(nevw + 1) * 16, o_be = o_al + (size_t)msw * 16,
                 o_x = o_be + (size_t)msw * 16, o_info = o_x + (size_t)msw * 16,
                 o_prob = (o_info + 256 + 255) / 256 * 256;
    int *hinfo = (int *)d->hqzd;
    bbjd_qz_problem *hp = (bbjd_qz_problem *)(d->hqzd + 256);
    zcx *hal = (zcx *)(d->hqzd + 256 + (sizeof(bbjd_qz_problem) + 255) / 256 * 256), *hbe = hal + msw, *hx = hbe + msw;
    zcx *stg = d->hgram;
    memcpy(stg, MA, cb);
    memcpy((char *)stg + cb, MB, cb);
    if (nconv > 0) memcpy(d->hsmall, conv, (size_t)nconv * 16);
    memset(hp, 0, sizeof(*hp));
    hp->k = (int)k; hp->ldm = (int)ldm;
    hp->MA = (const bbjd_zomplex *)db; hp->MB = (const bbjd_zomplex *)(db + mb);
    hp->nu0_re = creal(c->nu0); hp->nu0_im = cimag(c->nu0); hp->mu0_re = creal(c->mu0); hp->mu0_im = cimag(c->mu0);
    hp->sigma_re = creal(c->P->sigma); hp->sigma_im = cimag(c->P->sigma);
    hp->target = (int)c->P->target; hp->nkeep = (int)nkeep; hp->nconv = (int)nconv; hp->keep = (int)keep;
    hp->conv = (const bbjd_zomplex *)(db + o_conv);
    hp->tol = c->P->tol;
    hp->alpha = (bbjd_zomplex *)(db + o_al); hp->beta = (bbjd_zomplex *)(db + o_be); hp->x = (bbjd_zomplex *)(db + o_x);
    hp->zr = keep > 0 ? (bbjd_zomplex *)d->zr : NULL;
    hp->info = (int *)(db + o_info);
    int rc = 0;
    rc |= !HIPOK(hipMemcpyAsync(db, stg, cb, hipMemcpyHostToDevice, c->s));
    rc |= !HIPOK(hipMemcpyAsync(db + mb, (char *)stg + cb, cb, hipMemcpyHostToDevice, c->s));
    if (nconv > 0) rc |= !HIPOK(hipMemcpyAsync(db + o_conv, d->hsmall, (size_t)nconv * 16, hipMemcpyHostToDevice,
        c->s));
    rc |= !HIPOK(hipMemcpyAsync(db + o_prob, hp, sizeof(*hp), hipMemcpyHostToDevice, c->s));
    if (!rc) rc |= bbjd_qz_device_batch((const bbjd_qz_problem *)(db + o_prob), 1, c->s) ? 1 : 0;
    rc |= !HIPOK(hipMemcpyAsync(hinfo, db + o_info, 4 * sizeof(int), hipMemcpyDeviceToHost, c->s));
    rc |= !HIPOK(hipMemcpyAsync(hal, db + o_al, 3 * (size_t)msw * 16, hipMemcpyDeviceToHost, c->s));
    rc |= !HIPOK(dsync(c->d, c->s));
    if (rc) return -1;
    d->st.qz_device_steps++;
    if (hinfo[0] != BBJD_QZD_OK || hinfo[1] < 0 || hinfo[1] >= k) {
        d->st.qz_device_fallbacks++;
        d->st.qz_device_accuracy += hinfo[0] == BBJD_QZD_ACCURACY;
        return 1;
    }
    *p = hinfo[1];
    *ta = hal[*p];
    *tb = hbe[*p];
    memcpy(d->xv, hx, (size_t)k * sizeof(zcx));
    d->zr_ready = keep > 0;
    return 0;
}

static bbjd_status solve_body(const bbjd_params *P, bbjd_workspace *ws, bbjd_result *R, bbjd_int ms_solve)
{
    struct bbjd_dev *d = ws->dev;
    const double t_start = now_s();
    memset(&d->st, 0, sizeof(d->st));
    graphs_drop(d);
    /* n-sized arrays are laid out for this problem (leading dimension
     * roundup(n, 64)) inside the workspace's buffers, which may be sized for a
     * larger one: a resident workspace serving a small request then touches
     * only the start of each buffer instead of one stretch per column. */
    ctx_t cx = {ws, d, P, d->stream, P->n, (P->n + 63) / 64 * 64, ms_solve, ws->maxsub};
    ctx_t *c = &cx;
    const bbjd_int n = P->n, nev = P->nev, ms = ms_solve, ldm = ws->maxsub;
    d->nconv_last = 0;
    bbjd_int minsub = P->minsub > 0 ? P->minsub : (2 * nev + 5 < 20 ? 20 : 2 * nev + 5);
    const zcx sigma = P->sigma;
    const double tol = P->tol;
    zcx *MA = (zcx *)ws->MA, *MB = (zcx *)ws->MB;
    zcx *al = d->al, *be = d->be;
    bbjd_status st = BBJD_SUCCESS;
    int rc = 0;

    R->eigenvalues = calloc((size_t)nev, sizeof(zcx));
    R->eigenvectors = calloc((size_t)n * nev, sizeof(zcx));
    R->residuals = calloc((size_t)nev, sizeof(double));
    if (!R->eigenvalues || !R->eigenvectors || !R->residuals) {
        bbjd_result_free(R);
        R->status = BBJD_ERR_ALLOC;
        return BBJD_ERR_ALLOC;
    }
    zcx *eval = (zcx *)R->eigenvalues;

    zcx nu0 = 0.0, mu0 = 1.0;                                      /* :783-789 */
    if (P->target == BBJD_TARGET_NEAREST) {
        double sc = 1.0 / sqrt(1.0 + cabs(sigma) * cabs(sigma));
        nu0 = sc; mu0 = -sigma * sc;
    } else if (P->target == BBJD_TARGET_SMALLEST_MAG) {
        nu0 = 1.0; mu0 = 0.0;
    }
    cx.nu0 = nu0;
    cx.mu0 = mu0;
    ws->kdim = 0; ws->nlock = 0;
    d->pp_valid = 0;
    bbjd_qz_invalidate(&d->qz);
    rng_seed(d, 42);
    hipMemsetAsync(d->rd.flag, 0, FL_COUNT * 4, c->s);

    /* ---- start vector (:790-820) ---- */
    {
        zcx *hv = d->hx ? d->hx : malloc((size_t)n * sizeof(zcx));
        if (!hv) { st = BBJD_ERR_ALLOC; goto done; }
        if (P->v0) memcpy(hv, P->v0, (size_t)n * sizeof(zcx));
        else
            for (bbjd_int i = 0; i < n; i++) {
                double re = rng_unit(d) - 0.5;
                double im = rng_unit(d) - 0.5;
                hv[i] = re + im * I;
            }
        rc |= !HIPOK(hipMemcpyAsync(ws->V, hv, (size_t)n * 16, hipMemcpyHostToDevice, c->s));
        rc |= bbjd_k_norm(c->s, n, ws->V, &d->rd, SL_T0);
        rc |= sync_scal(c);
        if (!d->hx) free(hv);
    }
    if (rc) { st = BBJD_ERR_ALLOC; goto done; }
    if (d->hscal[SL_T0] < 1e-300) { st = BBJD_ERR_BREAKDOWN; goto done; }
    rc |= bbjd_k_scale(c->s, n, ws->V, 1.0 / d->hscal[SL_T0], 0.0);
    if (apply_ab(c, ws->V, ws->AV, ws->BV)) goto cb_err;
    rc |= bbjd_k_lincomb_norm(c->s, n, creal(nu0), cimag(nu0), ws->AV, creal(mu0), cimag(mu0), ws->BV, ws->W,
                              &d->rd, SL_W);
    rc |= bbjd_k_scale_inv(c->s, n, ws->W, d->rd.scal, SL_W, 0);
    if (rc || pencil_update(c, 0, 1, 0)) { st = BBJD_ERR_ALLOC; goto done; }
    ws->kdim = 1;

    for (bbjd_int iter = 0; iter < P->maxiter; iter++) {
        bbjd_int k = ws->kdim;
        const bbjd_int locked = ws->nlock;
        bbjd_int p = locked;
        d->st.outer_iters++;
        d->st.sum_k += k;
        /* ---- steps 1-2: projected QZ on the host, Ritz coefficients ---- */
        /* the ordered Schur form where a thick restart may follow (k >= ms),
         * else the eigenvalues and the target Ritz vector only */
        int schur = d->opt.qz_mode == BBJD_QZ_SCHUR || k >= ms;
        zcx ta = 0.0, tb = 1.0;
        d->zr_ready = 0;
        {
            const double tq = now_s();
            int qrc = 0;
            int on_device = 0;
            if (d->opt.qz_mode == BBJD_QZ_DEVICE) {
                bbjd_int nkeep = minsub - locked, keep = 0;
                if (nkeep < 1) nkeep = 1;
                if (k >= ms) {
                    keep = minsub < k ? minsub : k - 1;
                    if (keep < locked + 1) keep = locked + 1;
                }
                const int q = qz_device(c, k, MA, MB, eval, R->nconv, nkeep, keep, &p, &ta, &tb);
                if (q < 0) { st = BBJD_ERR_ALLOC; goto done; }
                on_device = q == 0;
                if (on_device) schur = 0;
            }
            if (!on_device && !schur) {
                qrc = bbjd_qz_target(&d->qz, k, MA, MB, nu0, mu0, al, be, sigma, P->target, eval, R->nconv, tol,
                                     &p);
                if (qrc == 2) {             /* ambiguous target: the reference's sequence decides */
                    schur = 1;
                    qrc = 0;
                    d->st.qz_schur_ties++;
                } else if (!qrc) {
                    ta = al[p];
                    tb = be[p];
                    qrc = bbjd_ritz_vector(&d->qz, k, MA, MB, ta, tb, d->xv);
                }
            }
            if (!on_device && schur) {
                bbjd_int nkeep = minsub - locked;
                if (nkeep < 1) nkeep = 1;
                /* a restart (RITZ mode, or DEVICE mode where the device step
                 * did not apply): the ordered Schur form of the harmonic
                 * standard problem (half the generalized QZ's cost), unless it
                 * defers a choice to the reference's sequence */
                int harm = 1;
                if (d->opt.qz_mode != BBJD_QZ_SCHUR && k >= ms && !c->w_mixed) {
                    harm = bbjd_qz_schur_harmonic(&d->qz, k, MA, MB, nu0, mu0, (zcx *)ws->ZB, al, be, sigma,
                                                  P->target, nkeep, eval, R->nconv, tol, &p, d->xv);
                    if (harm < 0) qrc = harm;
                    else if (harm == 0) {
                        ta = al[p];
                        tb = be[p];
                        d->st.qz_harmonic_restarts++;
                    }
                }
                if (harm > 0) {
                    qrc = bbjd_qz_solve(&d->qz, k, MA, MB, (zcx *)ws->SA, (zcx *)ws->SB, (zcx *)ws->ZA,
                                        (zcx *)ws->ZB, al, be, sigma, P->target, nkeep, eval, R->nconv, tol, &p);
                    if (!qrc) {
                        ta = al[p];
                        tb = be[p];
                        bbjd_ritz_coefficients(k, ldm, p, ta, tb, (const zcx *)ws->SA, (const zcx *)ws->SB,
                                               (const zcx *)ws->ZB, d->yv, d->xv);
                    }
                }
            }
            d->st.qz_seconds += now_s() - tq;
            if (qrc) { st = BBJD_ERR_LAPACK; goto done; }
        }
        /* ---- step 2: Ritz vector u = V x / ||V x|| on the device ---- */
        /* V is orthonormal, so ||V x|| = ||x||: x is normalised on the host
         * and u needs no norm pass and no scaling pass */
        void *u = ws->tmp1, *Au = ws->tmp2, *Bu = ws->tmp3, *r = ws->tmp4;
        double xnrm = 0;
        for (bbjd_int i = 0; i < k; i++) xnrm += creal(d->xv[i]) * creal(d->xv[i]) + cimag(d->xv[i]) * cimag(d->xv[i]);
        xnrm = sqrt(xnrm);
        if (!(xnrm > 1e-300)) { st = BBJD_ERR_BREAKDOWN; goto done; }
        for (bbjd_int i = 0; i < k; i++) d->xv[i] /= xnrm;
        if (k <= 128) rc |= bbjd_k_gemv_n_host(c->s, n, (int)k, ws->V, c->ld, d->xv, u);
        else {
            rc |= upload_small(c, d->rd.coef, d->xv, k);
            rc |= bbjd_k_gemv_n(c->s, n, (int)k, ws->V, c->ld, d->rd.coef, u, &d->rd, -1);
        }
        /* ---- step 3: residual ---- */
        if (d->opt.matvec_ABr && d->opt.callback_space == BBJD_CALLBACK_DEVICE) {
            hipStream_t prev = tl_cb_stream;
            tl_cb_stream = c->s;
            const int e = d->opt.matvec_ABr(n, (const bbjd_zomplex *)u, (bbjd_zomplex *)Au, (bbjd_zomplex *)Bu, tb, -ta,
                                            (bbjd_zomplex *)r, d->rd.scal + SL_R, P->op_data);
            tl_cb_stream = prev;
            if (e) goto cb_err;
        } else {
            if (apply_ab(c, u, Au, Bu)) goto cb_err;
            rc |= bbjd_k_lincomb_norm(c->s, n, creal(tb), cimag(tb), Au, -creal(ta), -cimag(ta), Bu, r, &d->rd,
                                      SL_R);
        }
        if (rc || sync_scal(c)) { st = BBJD_ERR_ALLOC; goto done; }
        d->st.stream_bytes += 16.0 * n * (k + 1 + 3);
        const double den = cabs(ta) + cabs(tb);
        const double rel = d->hscal[SL_R] / (den > 1e-300 ? den : 1.0);
        const zcx lam = cabs(tb) > 1e-300 ? ta / tb : ta;
        if (P->verbose >= 2)
            fprintf(stderr, "  iter %3lld  kdim %3lld  locked %2lld  eig (%+.6e, %+.6e)  res %.3e\n",
                    (long long)iter, (long long)k, (long long)locked, creal(lam), cimag(lam), rel);

        /* ---- step 4: convergence / locking (:905-932) ---- */
        if (rel < tol) {
            eval[R->nconv] = lam;
            rc |= !HIPOK(hipMemcpyAsync((zcx *)R->eigenvectors + (size_t)R->nconv * n, u, (size_t)n * 16,
                                        hipMemcpyDeviceToHost, c->s));
            rc |= !HIPOK(dsync(c->d, c->s));
            R->residuals[R->nconv] = rel;
            /* ||A u - lambda B u|| with ||u|| = 1: the FP64 residual of the
             * returned pair from the operator applies, r / theta_b */
            d->pair_res[R->nconv] = cabs(tb) > 1e-300 ? d->hscal[SL_R] / cabs(tb) : d->hscal[SL_R];
            R->nconv++;
            d->nconv_last = R->nconv;
            d->st.lock_iters++;
            if (P->verbose >= 1)
                fprintf(stderr, "  *** converged #%lld: (%+.10e, %+.10e)  res %.3e  iter %lld\n",
                        (long long)R->nconv, creal(lam), cimag(lam), rel, (long long)iter);
            if (R->nconv >= nev) { R->iters = iter + 1; goto done; }
            if (deflation_append(c, ws->nlock, u, Au, Bu, nu0, mu0)) { st = BBJD_ERR_ALLOC; goto done; }
            ws->nlock++;
            continue;          /* (MA, MB) unchanged: the QZ cache stays valid */
        }

        /* ---- step 6: correction (:944-972) ---- */
        /* Without a restart this iteration the correction is formed in place
         * in the next basis column (the expansion's copy is then skipped). */
        void *zc = ws->tmp2;
        if (P->inner_maxiter <= 0 && k < ms) zc = COLZ(ws->V, k);
        if (P->inner_maxiter > 0) {
            if (deflation_append(c, locked, u, Au, Bu, nu0, mu0)) { st = BBJD_ERR_ALLOC; goto done; }
            int g = gmres(c, ws->nlock + 1, ws->Q_right, ws->Q_left, ta, tb, r, zc, P->inner_maxiter,
                          P->inner_tol, 1);
            if (g == -1) { st = BBJD_ERR_CALLBACK; goto done; }
            if (g) { st = BBJD_ERR_ALLOC; goto done; }
        } else {
            void *sv = ws->tmp3;       /* Bu is not needed any more */
            if (P->precond && d->opt.precond2 && d->opt.callback_space == BBJD_CALLBACK_DEVICE) {
                /* z = M^-1 r and s = M^-1 u in one pass */
                hipStream_t prev = tl_cb_stream;
                tl_cb_stream = c->s;
                int e = d->opt.precond2(n, (const bbjd_zomplex *)r, (bbjd_zomplex *)zc,
                                        (const bbjd_zomplex *)u, (bbjd_zomplex *)sv, P->op_data);
                tl_cb_stream = prev;
                if (e) { st = BBJD_ERR_CALLBACK; goto done; }
            } else if (P->precond) {
                if (apply_op(c, (bbjd_matvec_fn)P->precond, r, zc)) { st = BBJD_ERR_CALLBACK; goto done; }
                if (apply_op(c, (bbjd_matvec_fn)P->precond, u, sv)) { st = BBJD_ERR_CALLBACK; goto done; }
            } else {
                rc |= !HIPOK(hipMemcpyAsync(zc, r, (size_t)n * 16, hipMemcpyDeviceToDevice, c->s));
                rc |= !HIPOK(hipMemcpyAsync(sv, u, (size_t)n * 16, hipMemcpyDeviceToDevice, c->s));
            }
            rc |= bbjd_k_svdv(c->s, n, u, zc, sv, &d->rd);
            d->st.stream_bytes += 16.0 * n * 6;
        }

        /* ---- step 7: thick restart (:975-1030) ---- */
        if (k >= ms) {
            bbjd_int keep = minsub < k ? minsub : k - 1;
            if (keep < locked + 1) keep = locked + 1;
            if (thick_restart(c, k, keep, nu0, mu0)) { st = BBJD_ERR_ALLOC; goto done; }
            ws->kdim = keep;
            k = keep;
        }

        /* ---- step 8: expansion (:1032-1068) ---- */
        {
            int e = expand(c, k, zc, nu0, mu0);
            if (e == -1) goto cb_err;
            if (e == -3) { R->iters = iter + 1; st = BBJD_ERR_BREAKDOWN; goto done; }
            if (e) { st = BBJD_ERR_ALLOC; goto done; }
        }
        bbjd_qz_invalidate(&d->qz);
        ws->kdim = k + 1;
        R->iters = iter + 1;
    }
    if (R->nconv < nev) st = BBJD_MAX_ITER;
done:
    {
        int flags[FL_COUNT];
        if (HIPOK(hipMemcpyAsync(flags, d->rd.flag, sizeof(flags), hipMemcpyDeviceToHost, c->s)) &&
            HIPOK(dsync(c->d, c->s)))
            d->st.dgks_second_passes = flags[FL_DGKS_COUNT];
        /* second CGS passes at the mean subspace size: one more read of the
         * basis in the fused form (k <= 128), two otherwise */
        if (d->st.outer_iters > 0) {
            const double kbar = (double)d->st.sum_k / (double)d->st.outer_iters;
            d->st.stream_bytes += 16.0 * n * d->st.dgks_second_passes * (kbar <= 128 ? kbar + 2 : 2 * kbar + 3);
        }
    }
    d->st.solve_seconds = now_s() - t_start;
    R->status = st;
    return st;
cb_err:
    dsync(c->d, c->s);
    d->st.solve_seconds = now_s() - t_start;
    R->status = BBJD_ERR_CALLBACK;
    return BBJD_ERR_CALLBACK;
}

/* a solve counts as active for its workspace's batcher, which then knows
 * whether more inner solves may still come */
static bbjd_status solve_impl(const bbjd_params *P, bbjd_workspace *ws, bbjd_result *R, bbjd_int ms_solve)
{
    bbjd_batcher *B = ws->dev ? (bbjd_batcher *)ws->dev->opt.inner_batcher : NULL;
    if (B) batcher_enter(B);
    const bbjd_status st = solve_body(P, ws, R, ms_solve);
    if (B) batcher_leave(B);
    return st;
}

bbjd_status bbjd_solve(const bbjd_params *P, bbjd_workspace *ws, bbjd_result *R)
{
    if (!P || !ws || !R) return BBJD_ERR_INVALID;
    memset(R, 0, sizeof(*R));
    /* reference semantics: the workspace was allocated for this problem and
     * its maxsub is the solve's (reference src/jd_core.c:749) */
    if (P->n != ws->n || P->nev > ws->nev || P->nev <= 0 || !P->matvec_A || !P->matvec_B ||
        (P->inner_maxiter > 0 && P->inner_maxiter > ws->inner_maxiter)) {
        R->status = BBJD_ERR_INVALID;
        return BBJD_ERR_INVALID;
    }
    return solve_impl(P, ws, R, ws->maxsub);
}

bbjd_status bbjd_solve_in(const bbjd_params *P, bbjd_workspace *ws, bbjd_result *R)
{
    if (!P || !ws || !R) return BBJD_ERR_INVALID;
    memset(R, 0, sizeof(*R));
    const bbjd_int ms = auto_maxsub(P);
    if (P->n <= 0 || P->n > ws->n || P->nev > ws->nev || P->nev <= 0 || ms > ws->maxsub || ms < 2 ||
        !P->matvec_A || !P->matvec_B || (P->inner_maxiter > 0 && P->inner_maxiter > ws->inner_maxiter)) {
        R->status = BBJD_ERR_INVALID;
        return BBJD_ERR_INVALID;
    }
    return solve_impl(P, ws, R, ms);
}

int bbjd_workspace_pair_residuals(const bbjd_workspace *ws, double *out)
{
    if (!ws || !ws->dev || !out) return -1;
    memcpy(out, ws->dev->pair_res, (size_t)ws->dev->nconv_last * sizeof(double));
    return (int)ws->dev->nconv_last;
}

/* ------------------------------------------------------------------ */
/* reference test helpers on host arrays                               */
/* ------------------------------------------------------------------ */
typedef struct {
    hipStream_t s;
    void *blk;
    bbjd_red rd;
    double *hscal;
} helper_t;

static int helper_open(helper_t *h, bbjd_int n, bbjd_int k, size_t extra, void **extra_ptr)
{
    memset(h, 0, sizeof(*h));
    if (!HIPOK(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking))) return -1;
    const int64_t pe = bbjd_red_part_entries(n, (int)(k + 2));
    size_t off_part = 0, off_pn = off_part + (size_t)pe * 16;
    size_t off_h = off_pn + ((size_t)(n + BBJD_TILE_MIN_ROWS - 1) / BBJD_TILE_MIN_ROWS + 64) * 8;
    size_t kk = (size_t)(k + 8) * 16;
    size_t off_h2 = off_h + kk, off_hs = off_h2 + kk, off_sc = off_hs + kk, off_fl = off_sc + 256,
           off_cn = off_fl + 64, off_co = off_cn + BBJD_CNT_SLOTS * 4, off_c2 = off_co + kk,
           off_ex = (off_c2 + 64 * 16 + 255) / 256 * 256;
    if (!HIPOK(hipMalloc(&h->blk, off_ex + extra))) { h->blk = NULL; return -1; }
    char *b = (char *)h->blk;
    h->rd.part = b + off_part; h->rd.pnorm = (double *)(b + off_pn); h->rd.h = b + off_h; h->rd.h2 = b + off_h2;
    h->rd.hsum = b + off_hs; h->rd.scal = (double *)(b + off_sc); h->rd.flag = (int *)(b + off_fl);
    h->rd.coef = b + off_co; h->rd.coef2 = b + off_c2; h->rd.part_cap = pe;
    h->rd.cnt = (int *)(b + off_cn);
    *extra_ptr = b + off_ex;
    if (!HIPOK(hipMemsetAsync(h->rd.flag, 0, 64 + BBJD_CNT_SLOTS * 4, h->s))) return -1;
    if (!HIPOK(hipHostMalloc((void **)&h->hscal, 256, 0))) { h->hscal = NULL; return -1; }
    return 0;
}

static void helper_close(helper_t *h)
{
    if (h->blk) hipFree(h->blk);
    if (h->hscal) hipHostFree(h->hscal);
    if (h->s) hipStreamDestroy(h->s);
}

double bbjd_block_orthogonalize(bbjd_int n, bbjd_int k, const bbjd_zomplex *Q, bbjd_zomplex *v,
                                bbjd_zomplex *h_work)
{
    helper_t h;
    void *dq = NULL;
    const size_t qb = (size_t)n * (k > 0 ? k : 0) * 16, vb = (size_t)n * 16;
    double res = NAN;
    if (helper_open(&h, n, k > 0 ? k : 1, qb + vb + 256, &dq) == 0) {
        void *dv = (char *)dq + (qb + 255) / 256 * 256;
        int rc = 0;
        if (k > 0) rc |= !HIPOK(hipMemcpyAsync(dq, Q, qb, hipMemcpyHostToDevice, h.s));
        rc |= !HIPOK(hipMemcpyAsync(dv, v, vb, hipMemcpyHostToDevice, h.s));
        rc |= bbjd_k_cgs2(h.s, n, (int)k, dq, n, dv, &h.rd, SL_T0, h.rd.hsum);
        rc |= !HIPOK(hipMemcpyAsync(v, dv, vb, hipMemcpyDeviceToHost, h.s));
        rc |= !HIPOK(hipMemcpyAsync(h.hscal, h.rd.scal, SL_COUNT * 8, hipMemcpyDeviceToHost, h.s));
        if (h_work && k > 0) rc |= !HIPOK(hipMemcpyAsync(h_work, h.rd.hsum, (size_t)k * 16, hipMemcpyDeviceToHost,
            h.s));
        rc |= !HIPOK(hipStreamSynchronize(h.s));
        if (!rc) res = h.hscal[SL_T0];
    }
    helper_close(&h);
    return res;
}

/* CGS2 + normalisation (unit = 1) or the two-pass projection (unit = 0) of
 * a host vector through the device kernels. */
static double host_orth_step(bbjd_int n, bbjd_int k, const bbjd_zomplex *Q, bbjd_zomplex *v, int unit, int *flags)
{
    helper_t h;
    void *dq = NULL;
    const size_t qb = (size_t)n * (k > 0 ? k : 0) * 16, vb = (size_t)n * 16;
    double res = NAN;
    if (helper_open(&h, n, k > 0 ? k : 1, qb + vb + 256, &dq) == 0) {
        void *dv = (char *)dq + (qb + 255) / 256 * 256;
        int rc = 0;
        if (k > 0) rc |= !HIPOK(hipMemcpyAsync(dq, Q, qb, hipMemcpyHostToDevice, h.s));
        rc |= !HIPOK(hipMemcpyAsync(dv, v, vb, hipMemcpyHostToDevice, h.s));
        if (unit) rc |= bbjd_k_cgs2_unit(h.s, n, (int)k, dq, n, dv, &h.rd, SL_T0);
        else rc |= bbjd_k_project_out(h.s, n, (int)k, dq, n, dv, &h.rd);
        rc |= !HIPOK(hipMemcpyAsync(v, dv, vb, hipMemcpyDeviceToHost, h.s));
        rc |= !HIPOK(hipMemcpyAsync(h.hscal, h.rd.scal, SL_COUNT * 8, hipMemcpyDeviceToHost, h.s));
        if (flags) rc |= !HIPOK(hipMemcpyAsync(flags, h.rd.flag, FL_COUNT * 4, hipMemcpyDeviceToHost, h.s));
        rc |= !HIPOK(hipStreamSynchronize(h.s));
        if (!rc) res = unit ? h.hscal[SL_T0] : 0.0;
    }
    helper_close(&h);
    return res;
}

double bbjd_block_orthonormalize(bbjd_int n, bbjd_int k, const bbjd_zomplex *Q, bbjd_zomplex *v, int *flags)
{
    return host_orth_step(n, k, Q, v, 1, flags);
}

int bbjd_project_out_host(bbjd_int n, bbjd_int k, const bbjd_zomplex *Q, bbjd_zomplex *v)
{
    return host_orth_step(n, k, Q, v, 0, NULL) == 0.0 ? 0 : -1;
}

double bbjd_expand_test_space(bbjd_int n, bbjd_int k, const bbjd_zomplex *W, bbjd_zomplex ca, bbjd_zomplex cb,
                              const bbjd_zomplex *a, const bbjd_zomplex *b, bbjd_zomplex *w, bbjd_zomplex *colA,
                              bbjd_zomplex *colB)
{
    helper_t h;
    void *dW = NULL;
    const size_t wb = (size_t)n * (k > 0 ? k : 0) * 16, vb = ((size_t)n * 16 + 255) / 256 * 256,
                 colb = ((size_t)(k + 1) * 16 + 255) / 256 * 256;
    double res = NAN;
    if (k < 1) return NAN;
    if (helper_open(&h, n, k, (wb + 255) / 256 * 256 + 3 * vb + 2 * colb, &dW) == 0) {
        char *p = (char *)dW + (wb + 255) / 256 * 256;
        void *dw = p, *da = p + vb, *db = p + 2 * vb, *dA = p + 3 * vb, *dB = p + 3 * vb + colb;
        int rc = 0;
        rc |= !HIPOK(hipMemcpyAsync(dW, W, wb, hipMemcpyHostToDevice, h.s));
        rc |= !HIPOK(hipMemcpyAsync(da, a, (size_t)n * 16, hipMemcpyHostToDevice, h.s));
        rc |= !HIPOK(hipMemcpyAsync(db, b, (size_t)n * 16, hipMemcpyHostToDevice, h.s));
        int pc = rc ? -1
                    : bbjd_k_cgs2_unit_pencil(h.s, n, (int)k, dW, n, dw, da, db, creal(ca), cimag(ca), creal(cb),
                                              cimag(cb), dA, dB, &h.rd, SL_T0);
        if (pc > 0) {
            /* unfused shape: the expansion's other route */
            rc |= bbjd_k_lincomb(h.s, n, creal(ca), cimag(ca), da, creal(cb), cimag(cb), db, dw);
            rc |= bbjd_k_cgs2_unit(h.s, n, (int)k, dW, n, dw, &h.rd, SL_T0);
            rc |= bbjd_k_gemv_c(h.s, n, (int)k, dW, n, da, db, dA, dB, &h.rd);
            rc |= bbjd_k_gemv_c(h.s, n, 1, dw, n, da, db, (char *)dA + (size_t)k * 16, (char *)dB + (size_t)k * 16,
                                &h.rd);
        } else if (pc < 0) rc = 1;
        rc |= !HIPOK(hipMemcpyAsync(w, dw, (size_t)n * 16, hipMemcpyDeviceToHost, h.s));
        rc |= !HIPOK(hipMemcpyAsync(colA, dA, (size_t)(k + 1) * 16, hipMemcpyDeviceToHost, h.s));
        rc |= !HIPOK(hipMemcpyAsync(colB, dB, (size_t)(k + 1) * 16, hipMemcpyDeviceToHost, h.s));
        rc |= !HIPOK(hipMemcpyAsync(h.hscal, h.rd.scal, SL_COUNT * 8, hipMemcpyDeviceToHost, h.s));
        rc |= !HIPOK(hipStreamSynchronize(h.s));
        if (!rc) res = h.hscal[SL_T0];
    }
    helper_close(&h);
    return res;
}

int bbjd_block_rotate(bbjd_int n, bbjd_int k, bbjd_int keep, bbjd_zomplex *X, const bbjd_zomplex *Z)
{
    if (n <= 0 || k <= 0 || keep <= 0 || keep > k) return -1;
    helper_t h;
    void *dX = NULL;
    const size_t xb = (size_t)n * k * 16, zb = (size_t)k * keep * 16;
    int rc = -1;
    zcx *zr = malloc(zb);
    if (!zr) return -1;
    /* row-major Z for the rotation: zr[j keep + c] = Z[j, c] */
    for (bbjd_int j = 0; j < k; j++)
        for (bbjd_int c = 0; c < keep; c++) zr[j * keep + c] = ((const zcx *)Z)[j + c * k];
    if (helper_open(&h, n, k, (xb + 255) / 256 * 256 + zb, &dX) == 0) {
        void *dZ = (char *)dX + (xb + 255) / 256 * 256;
        int e = 0;
        e |= !HIPOK(hipMemcpyAsync(dX, X, xb, hipMemcpyHostToDevice, h.s));
        e |= !HIPOK(hipMemcpyAsync(dZ, zr, zb, hipMemcpyHostToDevice, h.s));
        e |= bbjd_k_rotate(h.s, n, (int)k, (int)keep, dX, n, dZ);
        e |= !HIPOK(hipMemcpyAsync(X, dX, xb, hipMemcpyDeviceToHost, h.s));
        e |= !HIPOK(hipStreamSynchronize(h.s));
        rc = e ? -1 : 0;
    }
    helper_close(&h);
    free(zr);
    return rc;
}

int bbjd_block_gram(bbjd_int n, bbjd_int p, const bbjd_zomplex *X, bbjd_int q, const bbjd_zomplex *Y,
                    bbjd_zomplex *G)
{
    if (n <= 0 || p <= 0 || q <= 0 || p > BBJD_MAXK || q > BBJD_MAXK) return -1;
    /* partial-sum room for at least one chunk of the p x q result */
    const bbjd_int nt = (n + BBJD_PROJ_MIN_ROWS - 1) / BBJD_PROJ_MIN_ROWS;
    bbjd_int kk = p > q ? p : q;
    const bbjd_int need = (p * q + 2 * nt - 1) / (2 * nt);
    if (need > kk) kk = need;
    helper_t h;
    void *dx = NULL;
    const size_t xb = (size_t)n * p * 16, yb = (size_t)n * q * 16, gb = (size_t)p * q * 16;
    const size_t oy = (xb + 255) / 256 * 256, og = oy + (yb + 255) / 256 * 256;
    int rc = -1;
    if (helper_open(&h, n, kk, og + gb, &dx) == 0) {
        void *dy = (char *)dx + oy, *dg = (char *)dx + og;
        rc = 0;
        rc |= !HIPOK(hipMemcpyAsync(dx, X, xb, hipMemcpyHostToDevice, h.s));
        rc |= !HIPOK(hipMemcpyAsync(dy, Y, yb, hipMemcpyHostToDevice, h.s));
        rc |= bbjd_k_gram(h.s, n, (int)p, dx, n, (int)q, dy, n, dg, p, &h.rd);
        rc |= !HIPOK(hipMemcpyAsync(G, dg, gb, hipMemcpyDeviceToHost, h.s));
        rc |= !HIPOK(hipStreamSynchronize(h.s));
        rc = rc ? -1 : 0;
    }
    helper_close(&h);
    return rc;
}

double bbjd_mgs_orthogonalize(bbjd_int n, bbjd_int k, const bbjd_zomplex *Q, bbjd_zomplex *v)
{
    return bbjd_block_orthogonalize(n, k, Q, v, NULL);
}

int bbjd_correction_gmres(bbjd_int n, const bbjd_zomplex *Q_right, const bbjd_zomplex *Q_left,
                          bbjd_zomplex theta_a, bbjd_zomplex theta_b, const bbjd_zomplex *r, bbjd_zomplex *z,
                          bbjd_matvec_fn matvec_A, bbjd_matvec_fn matvec_B, bbjd_precond_fn precond,
                          void *op_data, bbjd_int maxiter, double tol, bbjd_workspace *ws)
{
    if (!ws || n != ws->n || maxiter <= 0 || maxiter > ws->inner_maxiter) return -1;
    struct bbjd_dev *d = ws->dev;
    bbjd_params P;
    bbjd_params_init(&P);
    P.n = n; P.nev = ws->nev; P.matvec_A = matvec_A; P.matvec_B = matvec_B; P.precond = precond;
    P.op_data = op_data; P.inner_maxiter = maxiter; P.inner_tol = tol;
    ctx_t cx = {ws, d, &P, d->stream, n, ws->ld, ws->maxsub, ws->maxsub};
    const bbjd_int ndefl = ws->nlock + 1;
    const int devspace = d->opt.callback_space == BBJD_CALLBACK_DEVICE;
    int rc = 0;
    const void *qr = Q_right, *ql = Q_left, *dr = r;
    void *dz = z;
    if (!devspace) {
        for (bbjd_int j = 0; j < ndefl; j++) {
            rc |= !HIPOK(hipMemcpyAsync(COLZ_STATIC(ws->Q_right, j, ws->ld), (const zcx *)Q_right + j * n,
                                        (size_t)n * 16, hipMemcpyHostToDevice, cx.s));
            rc |= !HIPOK(hipMemcpyAsync(COLZ_STATIC(ws->Q_left, j, ws->ld), (const zcx *)Q_left + j * n,
                                        (size_t)n * 16, hipMemcpyHostToDevice, cx.s));
        }
        rc |= !HIPOK(hipMemcpyAsync(ws->tmp4, r, (size_t)n * 16, hipMemcpyHostToDevice, cx.s));
        qr = ws->Q_right; ql = ws->Q_left; dr = ws->tmp4; dz = ws->tmp2;
    }   /* device space: Q_right/Q_left have leading dimension ws->ld */
    if (d->opt.precision == BBJD_PREC_MIXED)
        for (bbjd_int j = 0; j < ndefl; j++) {
            rc |= bbjd_k_z2c(cx.s, n, (const zcx *)qr + j * ws->ld, (c64 *)d->qrc + j * ws->ld);
            rc |= bbjd_k_z2c(cx.s, n, (const zcx *)ql + j * ws->ld, (c64 *)d->qlc + j * ws->ld);
        }
    if (rc) return -1;
    int g = gmres(&cx, ndefl, qr, ql, theta_a, theta_b, dr, dz, maxiter, tol, 0);
    if (g) return -1;
    if (!devspace) {
        rc |= !HIPOK(hipMemcpyAsync(z, dz, (size_t)n * 16, hipMemcpyDeviceToHost, cx.s));
        rc |= !HIPOK(hipStreamSynchronize(cx.s));
    }
    return rc ? -1 : 0;
}
