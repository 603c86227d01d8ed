/*
 * bbjd.h -- JDQZ eigensolver for the pencil A x = lambda B x, MI355X engine.
 *
 * Drop-in C interface: every type, enum value, struct field order and
 * function below has the meaning of the same name in the reference header
 * (kyle-augustson/BBJD include/bbjd.h), so a caller written against that
 * header compiles and runs unchanged against libbbjd.so.  The engine behind
 * it keeps the search/test subspaces in GPU memory (HBM) and runs the
 * n-sized arithmetic as HIP kernels for gfx950; the small projected QZ step
 * runs in FP64 LAPACK on the host.
 *
 * Layout: every dense subspace matrix is column-major; tall matrices in the
 * workspace have leading dimension ld >= n (see bbjd_workspace.ld).
 *
 * GPU-specific extensions (device-pointer callbacks, mixed precision,
 * streams, footprint query) live in bbjd_gpu.h; this header alone is the
 * reference-compatible surface.
 */
#ifndef BBJD_H
#define BBJD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
#include <complex>
extern "C" {
typedef std::complex<double> bbjd_zomplex;
#else
#include <complex.h>
/* reference include/bbjd.h:64 -- C99 double complex, 16 bytes (re, im) */
typedef double complex bbjd_zomplex;
#endif

/* reference include/bbjd.h:65 */
typedef int64_t bbjd_int;

/* Status codes; numeric values fixed by reference include/bbjd.h:70-78
 * (bbjd.py:248-257 maps them to messages). */
typedef enum {
    BBJD_SUCCESS       =  0,
    BBJD_MAX_ITER      =  1,   /* fewer than nev pairs converged      */
    BBJD_ERR_ALLOC     = -1,
    BBJD_ERR_INVALID   = -2,
    BBJD_ERR_CALLBACK  = -3,
    BBJD_ERR_LAPACK    = -4,
    BBJD_ERR_BREAKDOWN = -5
} bbjd_status;

/* Target selection; values fixed by reference include/bbjd.h:83-91 and
 * the wrapper's table (bbjd.py:48-56). */
typedef enum {
    BBJD_TARGET_NEAREST = 0,       /* nearest to sigma   */
    BBJD_TARGET_LARGEST_MAG,       /* largest |lambda|   */
    BBJD_TARGET_SMALLEST_MAG,      /* smallest |lambda|  */
    BBJD_TARGET_LARGEST_REAL,      /* largest Re lambda  */
    BBJD_TARGET_SMALLEST_REAL,     /* smallest Re lambda */
    BBJD_TARGET_LARGEST_IMAG,      /* largest Im lambda  */
    BBJD_TARGET_SMALLEST_IMAG      /* smallest Im lambda */
} bbjd_target;

/* Operator callbacks (reference include/bbjd.h:101-109): y = Op x, length n,
 * return 0 on success and nonzero to abort the solve with
 * BBJD_ERR_CALLBACK.  x and y never alias and are valid only during the
 * call; y must be written completely.  With the default host callback
 * space x and y are host memory; with BBJD_CALLBACK_DEVICE (bbjd_gpu.h)
 * they are device pointers and the callback enqueues its work on the
 * stream returned by bbjd_callback_stream(). */
typedef int (*bbjd_matvec_fn)(bbjd_int n, const bbjd_zomplex *x,
                              bbjd_zomplex *y, void *data);
typedef int (*bbjd_precond_fn)(bbjd_int n, const bbjd_zomplex *x,
                               bbjd_zomplex *y, void *data);

/* Solver parameters: field order of reference include/bbjd.h:114-150. */
typedef struct {
    bbjd_int        n;
    bbjd_int        nev;            /* eigenpairs wanted                 */
    bbjd_zomplex    sigma;          /* target shift                      */
    bbjd_target     target;
    bbjd_int        maxsub;         /* 0: min(max(3 nev + 10, 40), n)    */
    bbjd_int        minsub;         /* 0: max(2 nev + 5, 20)             */
    double          tol;            /* outer residual tolerance          */
    bbjd_int        maxiter;        /* outer iterations                  */
    bbjd_int        inner_maxiter;  /* 0: single SvdV step; >0: GMRES(m) */
    double          inner_tol;      /* GMRES relative tolerance          */
    const bbjd_zomplex *v0;         /* host start vector or NULL         */
    bbjd_matvec_fn  matvec_A;
    bbjd_matvec_fn  matvec_B;
    bbjd_precond_fn precond;        /* NULL: identity                    */
    void           *op_data;        /* shared by the three callbacks     */
    int             verbose;        /* 0 silent, 1 per pair, 2 per iter  */
} bbjd_params;

/* Result: field order of reference include/bbjd.h:155-162.  Arrays are
 * host memory allocated by bbjd_solve and released by bbjd_result_free,
 * also after an error return. */
typedef struct {
    bbjd_int        nconv;
    bbjd_zomplex   *eigenvalues;    /* nconv, in convergence order       */
    bbjd_zomplex   *eigenvectors;   /* n x nconv column-major, ||x||=1   */
    double         *residuals;      /* nconv                             */
    bbjd_int        iters;
    bbjd_status     status;
} bbjd_result;

struct bbjd_dev;   /* engine-private device state (streams, scratch)    */

/* Workspace: the reference's fields in the reference's order
 * (include/bbjd.h:170-224).  In this engine the n-sized arrays are DEVICE
 * pointers (HBM) with leading dimension `ld`; the maxsub x maxsub pencil,
 * Schur factors and the GMRES Hessenberg data are host arrays, because the
 * QZ step runs on the host.  swap_nm is NULL: the thick restart rotates the
 * basis in place.  Fields after h_work are engine-only. */
typedef struct {
    bbjd_int      n, maxsub, nev;
    bbjd_zomplex *V, *W, *AV, *BV;          /* device, ld x maxsub        */
    bbjd_zomplex *MA, *MB;                  /* host, maxsub x maxsub      */
    bbjd_zomplex *SA, *SB, *ZA, *ZB;        /* host, maxsub x maxsub      */
    bbjd_int      nlock;
    bbjd_int      kdim;
    bbjd_zomplex *gmres_V;                  /* device, ld x (ims + 1)     */
    bbjd_zomplex *gmres_H;                  /* host, (ims + 1) x ims      */
    bbjd_zomplex *gmres_cs, *gmres_sn, *gmres_g;   /* host                */
    bbjd_int      inner_maxiter;            /* ims = max(inner_maxiter,1) */
    bbjd_zomplex *tmp1, *tmp2, *tmp3, *tmp4;       /* device, n each      */
    bbjd_zomplex *swap_nm;                  /* NULL (in-place restart)    */
    bbjd_zomplex *Q_right, *Q_left;         /* device, ld x (nev + 1)     */
    bbjd_zomplex *h_work;                   /* device, 2 maxsub + nev + 4 */
    /* ---- engine-only ---- */
    bbjd_int        ld;                     /* leading dim of tall arrays */
    struct bbjd_dev *dev;
} bbjd_workspace;

/* reference include/bbjd.h:232 -- defaults nev 6, tol 1e-10, maxiter 200,
 * inner_tol 1e-2, target NEAREST, everything else zero. */
void bbjd_params_init(bbjd_params *p);

/* reference include/bbjd.h:235-236.  NULL on any allocation failure (device
 * or host), with partial allocations released.  Host callbacks, FP64. */
bbjd_workspace *bbjd_workspace_alloc(const bbjd_params *p);
void            bbjd_workspace_free(bbjd_workspace *ws);

/* reference include/bbjd.h:239-241 */
bbjd_status bbjd_solve(const bbjd_params *params, bbjd_workspace *ws,
                       bbjd_result *result);

/* reference include/bbjd.h:244 -- NULL-safe, nulls the pointers. */
void bbjd_result_free(bbjd_result *result);

/* reference include/bbjd.h:254-256 (implemented as CGS2 with the DGKS
 * 0.7071 test, reference src/jd_core.c:207-266).  Q (n x k, ld n) and v are
 * HOST arrays; the projection runs on the GPU. Returns ||v|| afterwards. */
double bbjd_mgs_orthogonalize(bbjd_int n, bbjd_int k, const bbjd_zomplex *Q,
                              bbjd_zomplex *v);

/* reference src/jd_core.c:207 (exported there though undeclared): the same
 * CGS2 with caller scratch h_work (>= k entries, host). */
double bbjd_block_orthogonalize(bbjd_int n, bbjd_int k, const bbjd_zomplex *Q,
                                bbjd_zomplex *v, bbjd_zomplex *h_work);

/* reference include/bbjd.h:268-282 with the argument meaning of
 * src/jd_core.c:568-582: Q_right and Q_left are n x (ws->nlock + 1)
 * matrices (right projector first).  Left-preconditioned GMRES on
 *   M^-1 P_L (theta_b A - theta_a B) P_R t = -M^-1 P_L r.
 * Q_right, Q_left, r and z are in the workspace's callback space (host by
 * default).  Returns 0, or -1 when a callback fails. */
int bbjd_correction_gmres(bbjd_int n,
                          const bbjd_zomplex *Q_right,
                          const bbjd_zomplex *Q_left,
                          bbjd_zomplex theta_a, bbjd_zomplex theta_b,
                          const bbjd_zomplex *r, bbjd_zomplex *z,
                          bbjd_matvec_fn matvec_A, bbjd_matvec_fn matvec_B,
                          bbjd_precond_fn precond, void *op_data,
                          bbjd_int maxiter, double tol, bbjd_workspace *ws);

#ifdef __cplusplus
}
#endif
#endif /* BBJD_H */
