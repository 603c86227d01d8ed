/*
 * bbjd_gpu.h -- engine extensions beyond the reference C surface.
 *
 * None of these has a counterpart in the reference (kyle-augustson/BBJD);
 * they select how the MI355X engine runs a solve.  bbjd_workspace_alloc()
 * in bbjd.h is bbjd_workspace_alloc_gpu(p, NULL): host callbacks, FP64.
 */
#ifndef BBJD_GPU_H
#define BBJD_GPU_H

#include "bbjd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Where the callback vectors x, y live. */
typedef enum {
    BBJD_CALLBACK_HOST   = 0,  /* host memory; the engine stages D2H/H2D    */
    BBJD_CALLBACK_DEVICE = 1   /* device pointers, ordered on the engine's
                                  stream (bbjd_callback_stream())            */
} bbjd_callback_space;

/* Arithmetic of the correction solve.  The outer loop (orthogonalisation,
 * projected pencil, residuals, QZ) is FP64 in both. */
typedef enum {
    BBJD_PREC_FP64  = 0,       /* everything complex128                     */
    BBJD_PREC_MIXED = 1        /* inner GMRES + preconditioner in complex64,
                                  FP32 accumulation                          */
} bbjd_precision;

/* complex64 element, layout (re, im) of float */
typedef struct { float re, im; } bbjd_cfloat;

/* Two preconditioner applies at once, y0 = M^-1 x0 and y1 = M^-1 x1, in the
 * workspace's callback space; same return contract as bbjd_precond_fn. */
typedef int (*bbjd_precond2_fn)(bbjd_int n, const bbjd_zomplex *x0, bbjd_zomplex *y0,
                                const bbjd_zomplex *x1, bbjd_zomplex *y1, void *data);

/* complex64 device-pointer operator, same contract as bbjd_matvec_fn */
typedef int (*bbjd_op_c64_fn)(bbjd_int n, const bbjd_cfloat *x,
                              bbjd_cfloat *y, void *data);

/* Both matvecs of one vector at once, ya = A x and yb = B x (device
 * pointers, no aliasing), same return contract as bbjd_matvec_fn; and its
 * complex64 form. */
typedef int (*bbjd_matvec2_fn)(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *ya,
                               bbjd_zomplex *yb, void *data);
typedef int (*bbjd_op2_c64_fn)(bbjd_int n, const bbjd_cfloat *x, bbjd_cfloat *ya,
                               bbjd_cfloat *yb, void *data);

/* ya = A x, yb = B x and the residual r = alpha ya + beta yb with ||r||_2
 * written to the device double *rnorm, in one call (device pointers); same
 * return contract as bbjd_matvec_fn. */
typedef int (*bbjd_residual_fn)(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *ya, bbjd_zomplex *yb,
                                bbjd_zomplex alpha, bbjd_zomplex beta, bbjd_zomplex *r, double *rnorm,
                                void *data);

typedef struct {
    bbjd_callback_space callback_space;
    bbjd_precision      precision;
    void               *stream;      /* hipStream_t; NULL: workspace owns one */
    /* complex64 operators for BBJD_PREC_MIXED (device pointers).  NULL
     * entries fall back to the FP64 callbacks through conversion kernels. */
    bbjd_op_c64_fn      matvec_A_c64;
    bbjd_op_c64_fn      matvec_B_c64;
    bbjd_op_c64_fn      precond_c64;
    /* optional: the Sleijpen-van der Vorst step's two preconditioner applies
     * (M^-1 r, M^-1 u) as one call; NULL: two precond calls.  Must apply the
     * same M as params.precond. */
    bbjd_precond2_fn    precond2;
    /* projected step: BBJD_QZ_RITZ (default) computes the ordered
     * generalized Schur form only on iterations a thick restart may follow
     * and otherwise the eigenvalues alone plus the target Ritz vector;
     * BBJD_QZ_SCHUR computes the ordered Schur form every iteration, the
     * reference's sequence of LAPACK calls.  Same Ritz pairs in exact
     * arithmetic. */
    int                 qz_mode;
    /* 1: capture the inner GMRES loop's device work into HIP graphs, once
     * per deflation count (the prologue and each Arnoldi step), and replay
     * them; the kernels are the same as without, so are the results.  Takes
     * effect with BBJD_CALLBACK_DEVICE only, and the device callbacks must
     * then be capturable: kernels and copies enqueued on the callback
     * stream, no synchronisation, allocation or memset. */
    int                 graphs;
    /* optional, device callbacks only: A x and B x of the same vector in one
     * call (and its complex64 form for the mixed inner solve); must apply
     * the same A and B as params.matvec_A / matvec_B.  NULL: two calls. */
    bbjd_matvec2_fn     matvec_AB;
    bbjd_op2_c64_fn     matvec_AB_c64;
    /* optional, device callbacks only: the residual step's A u, B u and
     * r = theta_b A u - theta_a B u with ||r|| in one call (same A, B). */
    bbjd_residual_fn    matvec_ABr;
    /* inner solve of BBJD_PREC_MIXED with a preconditioner:
     * BBJD_INNER_REFERENCE (default) is the reference's GMRES on
     * M^-1 P_L (theta_b A - theta_a B) P_R, two fixed projections on each
     * side per step; BBJD_INNER_PROJECTED uses the projected preconditioner
     * of Fokkema, Sleijpen and van der Vorst (SIAM J. Sci. Comput. 20, 1998):
     * with Yh = M^-1 Q_left and H = Q_right^H Yh (both kept on the device
     * across outer iterations, H^-1 formed on the host), GMRES runs on
     * (I - Yh H^-1 Q_right^H) M^-1 (theta_b A - theta_a B) over Q_right-perp,
     * one projection and one update per step.  Same exact solution of the
     * correction equation; the outer FP64 loop and its convergence test are
     * unchanged.  Falls back to the reference form for a call whose H is
     * numerically singular. */
    int                 inner_form;
    /* 1: the engine's host waits on its stream sleep until the GPU signals
     * (events created with hipEventBlockingSync) instead of spinning, so a
     * thread waiting on the GPU leaves its core to others (many workspaces
     * solving at once on a busy host); 0: the process's default wait. */
    int                 blocking_sync;
    /* a bbjd_batcher shared by workspaces that solve at once (serving):
     * their inner solves of BBJD_PREC_MIXED in the reference form, with the
     * library's banded operators (bbjd_op_*_c64 as device callbacks), run
     * together on the batcher's threads, each Arnoldi step's kernels of all
     * the solves in the batch merged into shared launches; every solve's
     * arithmetic is the one it does alone (same kernels, grids and order),
     * so are its results.  Other inner solves, and NULL, run on their own. */
    void               *inner_batcher;
} bbjd_gpu_options;

enum { BBJD_INNER_REFERENCE = 0, BBJD_INNER_PROJECTED = 1 };

/* BBJD_QZ_DEVICE: the projected step on the GPU (bbjd_qz_device_batch, one
 * wavefront) for k <= 64 with the nearest-sigma or smallest-magnitude test
 * space, the host step (as BBJD_QZ_RITZ) where it does not apply or flags
 * the iteration. */
enum { BBJD_QZ_RITZ = 0, BBJD_QZ_SCHUR = 1, BBJD_QZ_DEVICE = 2 };

void bbjd_gpu_options_init(bbjd_gpu_options *o);

/* Inner-solve batcher: `threads` threads, each with its own HIP stream, run
 * batches of up to max_jobs (<= 16) inner solves submitted by workspaces
 * that name it in bbjd_gpu_options.inner_batcher.  A thread takes the solves
 * waiting when it is free, and when fewer than max_jobs are waiting while
 * other solves of the batcher's workspaces are still running, waits up to
 * window_us for more.  inner_maxiter: the largest of the workspaces'.  No
 * counterpart in the reference (one solve per process there). */
typedef struct bbjd_batcher bbjd_batcher;
bbjd_batcher *bbjd_batcher_create(int threads, int max_jobs, double window_us, bbjd_int inner_maxiter);
void          bbjd_batcher_destroy(bbjd_batcher *b);
typedef struct {
    long batches;       /* batches run                                   */
    long jobs;          /* inner solves run in them                      */
    long steps;         /* Arnoldi steps (one per solve per step)        */
    long recorded;      /* kernel launches and copies the solves made    */
    long launches;      /* launches that ran them                        */
} bbjd_batcher_stats;
void bbjd_batcher_get_stats(const bbjd_batcher *b, bbjd_batcher_stats *st);

/* Exact device bytes bbjd_workspace_alloc_gpu() will take for (p, o);
 * the serving loop admits requests against this figure. */
size_t bbjd_workspace_bytes(const bbjd_params *p, const bbjd_gpu_options *o);

bbjd_workspace *bbjd_workspace_alloc_gpu(const bbjd_params *p,
                                         const bbjd_gpu_options *o);

/* Solve on a resident workspace allocated for a larger shape: p->n,
 * p->nev, the solve's maxsub (p->maxsub, or the reference's auto rule
 * min(max(3 nev + 10, 40), n)) and p->inner_maxiter may each be at most the
 * workspace's.  bbjd_solve keeps the reference rule that the workspace's
 * maxsub is the solve's (reference src/jd_core.c:749). */
bbjd_status bbjd_solve_in(const bbjd_params *p, bbjd_workspace *ws, bbjd_result *r);

/* ||A x_j - lambda_j B x_j|| (||x_j|| = 1) of each pair converged in the last
 * solve on ws, from the FP64 operator applies of the convergence test.
 * Writes nconv values, returns nconv (or -1). */
int bbjd_workspace_pair_residuals(const bbjd_workspace *ws, double *out);

/* The stream on which the engine is running the current callback of the
 * calling thread (hipStream_t), for DEVICE-space callbacks. */
void *bbjd_callback_stream(void);

/* hipStream_t the workspace runs on */
void *bbjd_workspace_stream(const bbjd_workspace *ws);

/* Per-solve counters of the last bbjd_solve on a workspace. */
typedef struct {
    bbjd_int outer_iters;       /* iterations of the main loop            */
    bbjd_int lock_iters;        /* of which ended in a lock               */
    bbjd_int restarts;
    bbjd_int inner_steps;       /* GMRES Arnoldi steps, all solves        */
    bbjd_int dgks_second_passes;/* CGS2 calls that took the second pass   */
    bbjd_int sum_k;             /* sum of subspace sizes over iterations  */
    double   qz_seconds;        /* host ZGGES/ZTGSEN + selection          */
    double   solve_seconds;     /* wall time of bbjd_solve                */
    double   stream_bytes;      /* algorithmic HBM bytes of the n-sized
                                   engine kernels (operators excluded)    */
    bbjd_int restart_fallbacks; /* restarts whose Cholesky QR of the test
                                   space fell back to column-wise CGS2    */
    bbjd_int graph_captures;    /* inner-solve segments captured (graphs)  */
    bbjd_int graph_launches;    /* and launched, captures included         */
    bbjd_int qz_schur_ties;     /* iterations between restarts whose target
                                   was a tie up to rounding (or had no
                                   unconverged candidate), decided by the
                                   ordered Schur form as the reference does */
    bbjd_int qz_device_steps;   /* projected steps run on the device         */
    bbjd_int qz_device_fallbacks; /* of which flagged back to the host      */
    bbjd_int qz_device_accuracy;  /* of those, flagged because the values'
                                     error bounds could change a choice     */
    bbjd_int inner_steps_discarded; /* Arnoldi steps run ahead of the inner
                                       solve's stop and not used            */
    bbjd_int qz_harmonic_restarts;  /* restarts whose ordered Schur form came
                                       from the harmonic standard problem   */
    bbjd_int inner_projected;       /* inner solves in BBJD_INNER_PROJECTED  */
    bbjd_int inner_projected_fallbacks; /* of which took the reference form
                                       (H numerically singular)             */
} bbjd_stats;

int bbjd_workspace_stats(const bbjd_workspace *ws, bbjd_stats *st);

/* First GPU call of a process that serves many solves from many host
 * threads: sync_yield = 1 makes threads waiting on the GPU yield their core
 * (the host QZ of other solves needs it), 2 makes them block until the GPU
 * signals.  Returns 0 (also when the device was already initialised and the
 * flag could not change). */
int bbjd_device_init(int sync_yield);

/* Launch timing of the engine's dominant streaming kernel (the one-vector
 * projection Q^H x): enable, run, then harvest the number of timed launches,
 * the sum of their durations (ms, between start and stop HIP events carried
 * by each dispatch) and of their algorithmic bytes (16 n (k + 1): Q and x
 * read once).  First CGS passes and plain projections; the conditional
 * second passes are not timed. */
void bbjd_ktime_enable(int on);
int  bbjd_ktime_harvest(double *total_ms, double *total_bytes);

/* G (p x q, column-major) = X^H Y for host arrays X (n x p) and Y (n x q),
 * column-major with leading dimension n, through the engine's device Gram
 * kernel (the restart's test-space QR and projected pencil).  No
 * counterpart in the reference; a utility for host callers and tests.
 * Returns 0, or -1 on invalid sizes or a device error. */
int bbjd_block_gram(bbjd_int n, bbjd_int p, const bbjd_zomplex *X, bbjd_int q, const bbjd_zomplex *Y,
                    bbjd_zomplex *G);

/* Test entries on host arrays (no counterpart in the reference; the
 * engine's own steps, for parity tests against the oracle).
 * bbjd_block_orthonormalize: the expansion's CGS2 of v against Q (n x k,
 * leading dimension n) followed by v /= ||v|| when ||v|| > 1e-300, as the
 * reference's expansion does after bbjd_block_orthogonalize
 * (src/jd_core.c:1038-1058); returns ||v|| before the scaling, NAN on a
 * device error; flags (8 ints, may be NULL) receives the device decision
 * flags, flags[0] = 1 when the DGKS second pass ran.
 * bbjd_project_out_host: v -= Q (Q^H v), two fixed passes (reference
 * src/jd_core.c:269-309); returns 0 or -1. */
double bbjd_block_orthonormalize(bbjd_int n, bbjd_int k, const bbjd_zomplex *Q, bbjd_zomplex *v, int *flags);
/* bbjd_expand_test_space: the expansion's test-space step (reference
 * src/jd_core.c:1060-1104) with a = A v, b = B v: w = normalised CGS2 of
 * ca a + cb b against W (n x k, k >= 1), and the new pencil column
 * colA = [W w]^H a, colB = [W w]^H b (k + 1 entries each).  Returns ||w||
 * before the scaling, NAN on a device error. */
double bbjd_expand_test_space(bbjd_int n, bbjd_int k, const bbjd_zomplex *W, bbjd_zomplex ca, bbjd_zomplex cb,
                              const bbjd_zomplex *a, const bbjd_zomplex *b, bbjd_zomplex *w, bbjd_zomplex *colA,
                              bbjd_zomplex *colB);
int    bbjd_project_out_host(bbjd_int n, bbjd_int k, const bbjd_zomplex *Q, bbjd_zomplex *v);

/* Blocks per Gram launch (the restart's test-space QR and projected pencil):
 * row chunks x 64 x 64 output tiles, default 512.  For tuning. */
void bbjd_set_gram_blocks(int blocks);

/* The restart's in-place basis rotation: 1 (default) on the FP64 matrix
 * cores in the streaming form (input columns through LDS in chunks, any k,
 * keep <= 128; the staged form beyond); 2: the staged form (all k input
 * columns of a row tile in LDS, when they fit); 0: on vector FP64.  For
 * tests and tuning. */
void bbjd_set_rotate_mfma(int on);
/* Input columns per LDS chunk of the streaming rotation: 8, 16 or 32 (0:
 * chosen from the output width).  For tuning. */
void bbjd_set_rotate_chunk(int kc);

/* Arnoldi steps of the inner GMRES the host keeps in flight (1..8, default
 * 4; 1: one step at a time).  With device callbacks and no graphs a step
 * needs nothing from the host, so the host reads each step's Hessenberg
 * column while the next ones run; steps past the reference's stop are
 * discarded, with the same results as one at a time.  Process-wide. */
void bbjd_set_gmres_lookahead(int steps);

/* Test entry: X[:, :keep] = X[:, :k] Z in place for host arrays X (n x k,
 * leading dimension n) and Z (k x keep, column-major) through the engine's
 * restart rotation (reference src/jd_core.c:983-1001); returns 0 or -1. */
int bbjd_block_rotate(bbjd_int n, bbjd_int k, bbjd_int keep, bbjd_zomplex *X, const bbjd_zomplex *Z);

/* 1 (default): CGS2 and the two-pass projection take the three-read form
 * (first update and second projection in one pass over Q) for k <= 128;
 * 0: the four-launch form everywhere; 2: the three-read form with the
 * normalisation after a second pass always by the computed norm (the rare
 * route of 1, forced).  For tests and tuning. */
void bbjd_set_cgs_fused(int mode);

/* Kernel benchmark (tile autotuning and roofline checks): times one engine
 * kernel in isolation at shape (n, k) on a private stream, reps launches.
 * which: 0 projection Q^H x, 1 projection of two vectors, 2 update
 * x -= Q h, 3 in-place basis rotation k -> k/2, 4 full CGS2, 5 a plain copy
 * of n complex values (the reference for counter-measured traffic), 6 the
 * two-pass projection with both passes run.  Writes
 * the average ms per launch and the algorithmic bytes per launch. */
int bbjd_kbench(int which, int64_t n, int k, int reps, double *avg_ms, double *bytes);

/* Grid size (blocks) up to which a reduction's finishing step runs in the
 * last block of the producing kernel instead of a second launch. */
void bbjd_set_fuse_blocks(int max_blocks);

/* 1 (default): the Sleijpen-van der Vorst step's axpy launch also reduces
 * its two dots' partials (in the finishing step's order, so bit-identical),
 * one launch fewer per outer iteration; 0: a finishing launch of its own.
 * Tuning and tests. */
void bbjd_set_svdv_fused(int on);

/* Largest grid of the fused CGS2's middle pass (default 512 blocks). */
void bbjd_set_cgs_mid_blocks(int max_blocks);

/* Tiles of the streaming kernels, for tuning.  proj_blocks: blocks per
 * projection launch (rows are split into chunks to give about that many;
 * 0 keeps the current value).  upd_rows: rows per block of the update / norm
 * kernels, 256, 512 or 1024, or 0 to choose per launch from the shape (the
 * largest tile that still gives blocks_target blocks; blocks_target 0 keeps
 * the current target).  Returns -2 on an unsupported value. */
int bbjd_set_tiles(int proj_blocks, int upd_rows, int blocks_target);

/* Per-shape tile table, as written by the autotuner (bbjd_amd/autotune.py):
 * nn x nk classes, class (i, j) = shapes with n <= n_upper[i] and
 * k <= k_upper[j] (the last class in each direction takes larger shapes);
 * proj_blocks[i nk + j] = blocks per projection launch, upd_rows[i nk + j] =
 * rows per update block (0 for the shape rule; else 64, 128, 256, 512 or
 * 1024).  nn = 0 clears the table (global defaults).  Not while solves run.
 * Returns 0 or -2 for an invalid table (at most 8 classes each way). */
int bbjd_set_tile_table(int nn, const int64_t *n_upper, int nk, const int *k_upper, const int *proj_blocks,
                        const int *upd_rows);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc: bbjd_block_orthogonalize (CGS2 with the DGKS test,
 * reference src/jd_core.c:207-241) on device arrays, with the reduction
 * scratch in a file-scope device array of the library (about 2.2 MB on each
 * GPU it runs on, held until the process ends) instead of a per-call
 * allocation; it calls no allocating routine.
 * Q (n x k, column-major, leading dimension ldq >= n), v (n entries,
 * orthogonalised in place), h_work (k entries: the summed coefficients of
 * both passes, or NULL) and norm_out (one double: ||v|| after, or NULL) are
 * device pointers.  Everything is enqueued on stream (hipStream_t; NULL:
 * the null stream) and nothing waits.  Limits: n <= 2^23, k <= 1024.
 * Returns 0; -1 for sizes beyond the limits or a NULL array (nothing
 * enqueued); -2 on a device error (v undefined).  Every call on one GPU
 * shares the scratch: calls in flight at the same time on one GPU (from
 * other streams or threads) race; calls ordered on one stream do not. */
int bbjd_block_orthogonalize_noalloc(bbjd_int n, bbjd_int k, const bbjd_zomplex *Q, bbjd_int ldq, bbjd_zomplex *v,
                                     bbjd_zomplex *h_work, double *norm_out, void *stream);

/* The projected step on the device, batched (no counterpart in the
 * reference, whose solve_projected runs ZGGES/ZTGSEN on the host,
 * src/jd_core.c:319-462).  One 64-lane wavefront per problem, k <= 64, all
 * data in LDS: with C = nu0 MA + mu0 MB upper triangular (the engine's
 * harmonic test space), the ordered Schur form M = Z T Z^H of
 * M = C^-1 MB, the reference's selection (values within 100 tol of a
 * converged one, then the nkeep best by the target key) moved to the
 * leading positions in order, the target (best unconverged value over all
 * positions), alpha / beta = theta per position with
 * alpha = 1 - mu0 t_ii, beta = nu0 t_ii, and the Ritz coefficients
 * x = Z y of the target.  info[0]: BBJD_QZD_* status (anything but OK asks
 * the caller to take the host step), info[1]: target position, info[2]:
 * number selected, info[3]: QR iterations (the last three 0 when the step
 * stopped early: not triangular, accuracy, no convergence, unsupported). */
typedef struct {
    int                 k, ldm;         /* size; leading dimension of MA, MB  */
    const bbjd_zomplex *MA, *MB;        /* device, column-major               */
    double              nu0_re, nu0_im, mu0_re, mu0_im, sigma_re, sigma_im;
    int                 target, nkeep, nconv, keep;
    const bbjd_zomplex *conv;           /* device, nconv converged values     */
    double              tol;
    bbjd_zomplex       *alpha, *beta;   /* device, k each, or NULL            */
    bbjd_zomplex       *x;              /* device, k: Ritz coefficients       */
    bbjd_zomplex       *zr;             /* device, k x keep row-major
                                           (zr[j keep + c] = Z(j, c)) or NULL */
    bbjd_zomplex       *zfull, *tfull;  /* device, k x k column-major, or NULL */
    int                *info;           /* device, 4 ints                     */
} bbjd_qz_problem;

enum { BBJD_QZD_OK = 0, BBJD_QZD_AMBIGUOUS = 1, BBJD_QZD_NOT_TRIANGULAR = 2, BBJD_QZD_ACCURACY = 3,
       BBJD_QZD_NO_CONVERGENCE = 4, BBJD_QZD_UNSUPPORTED = 5 };

/* nprob problems (probs: a device array) in one launch on stream
 * (hipStream_t, NULL: the null stream).  Returns 0 or -1. */
int bbjd_qz_device_batch(const bbjd_qz_problem *probs, int nprob, void *stream);

/* One problem on host arrays through bbjd_qz_device_batch (tests and
 * tools): MA, MB k x k with leading dimension ldm; outputs host arrays of
 * k (alpha, beta, x) and k x k column-major (z, t; any may be NULL), info
 * 4 ints.  Returns 0 or -1 on a device error. */
int bbjd_qz_device_host(bbjd_int k, const bbjd_zomplex *MA, const bbjd_zomplex *MB, bbjd_int ldm,
                        bbjd_zomplex nu0, bbjd_zomplex mu0, bbjd_zomplex sigma, int target, bbjd_int nkeep,
                        const bbjd_zomplex *conv, bbjd_int nconv, double tol, bbjd_zomplex *alpha,
                        bbjd_zomplex *beta, bbjd_zomplex *x, bbjd_zomplex *z, bbjd_zomplex *t, int *info);

/* Device memory helpers for callers without another GPU runtime (tests,
 * tools): allocate / free / synchronise the device. */
void *bbjd_dev_alloc(size_t bytes);
void  bbjd_dev_free(void *p);
int   bbjd_dev_sync(void);

/* Average ms of reps applies of a DEVICE-space operator f(n, x, y, data) on
 * the null stream (HIP events). */
int bbjd_time_op(bbjd_matvec_fn f, bbjd_int n, const void *x, void *y, void *data, int reps, double *avg_ms);

/* Device pointer -> host vector download helper for tests/tools. */
int bbjd_memcpy_d2h(void *dst, const void *src, size_t bytes);
int bbjd_memcpy_h2d(void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* BBJD_GPU_H */
