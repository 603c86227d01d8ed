/*
 * bbjd_ops.h -- device operators for block-banded pencils with dense tau
 * rows (the pencils the reference is built for: block-banded interior plus
 * dense boundary-condition rows, reference include/bbjd.h:8-9).
 *
 * The reference leaves the operators to the caller (scipy CSC matvec,
 * bbjd.py:183-205; SuperLU/FPLUS preconditioner, tests/test_bbjd.py:59-73).
 * These are the engine's own DEVICE-space callbacks for that structure:
 * pass bbjd_op_matvec_A / _B / _precond as params.matvec_A / matvec_B /
 * precond and the bbjd_op* as params.op_data, with callback_space
 * BBJD_CALLBACK_DEVICE.
 *
 * Pencil layout (host arrays handed to bbjd_op_create):
 *   n rows, block size bs (n % bs == 0), nb = n / bs block rows,
 *   block half-width hw (1 block-tridiagonal, 2 block-pentadiagonal).
 *   band: nb x (2 hw + 1) blocks of bs x bs, block (i, d) at offset
 *         (i (2 hw + 1) + d) bs^2 holds A[i bs + r, (i + d - hw) bs + c] at
 *         [c bs + r] (column-major block); out-of-range blocks are ignored.
 *   tau:  ntau x n row-major, the last ntau rows of A (dense); the band
 *         entries of those rows are ignored.
 *   bdiag: n, B = diag(bdiag).
 */
#ifndef BBJD_OPS_H
#define BBJD_OPS_H

#include "bbjd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    bbjd_int            n;
    int                 bs, hw, ntau;
    const bbjd_zomplex *band;
    const bbjd_zomplex *tau;
    const bbjd_zomplex *bdiag;
} bbjd_pencil_desc;

/* Seeded generator of the benchmark pencils: every band block and tau entry
 * iid CN(0,1) (real and imaginary parts N(0, 1/2), Box-Muller over a
 * splitmix64 counter hash, so the value of an entry depends only on
 * (seed, entry) and not on threading), plus diag_shift on the diagonal of
 * the diagonal blocks; bdiag = 1 except 0 on the ntau tau rows.
 * Arrays are caller-allocated with the sizes above.  Returns 0. */
int bbjd_pencil_generate(bbjd_int n, int bs, int hw, int ntau, uint64_t seed,
                         double diag_shift, bbjd_zomplex *band,
                         bbjd_zomplex *tau, bbjd_zomplex *bdiag);

typedef enum {
    BBJD_PC_NONE         = 0,
    BBJD_PC_EXACT_LU     = 1,  /* block cyclic reduction of the banded part,
                                  tau rows by Schur complement              */
    BBJD_PC_BLOCK_JACOBI = 2   /* inverses of the bs x bs diagonal blocks   */
} bbjd_pc_kind;

typedef struct bbjd_op bbjd_op;

/* Uploads the pencil, builds M = A - sigma B's preconditioner of kind pc.
 * want_c64 also keeps complex64 copies for the mixed-precision inner solve.
 * stream: hipStream_t used for set-up and by default for applies (NULL:
 * the null stream).  Returns NULL on failure. */
bbjd_op *bbjd_op_create(const bbjd_pencil_desc *d, bbjd_zomplex sigma,
                        bbjd_pc_kind pc, int want_c64, void *stream);
void     bbjd_op_free(bbjd_op *op);
size_t   bbjd_op_bytes(const bbjd_op *op);   /* device bytes held */
bbjd_int bbjd_op_size(const bbjd_op *op);    /* n */
/* device bytes bbjd_op_create(d, ., pc, want_c64, .) will take */
size_t   bbjd_op_bytes_for(const bbjd_pencil_desc *d, int pc, int want_c64);

/* DEVICE-space callbacks (bbjd_matvec_fn / bbjd_precond_fn); data is the
 * bbjd_op*.  They run on bbjd_callback_stream() when called by the engine,
 * else on the op's stream. */
int bbjd_op_matvec_A(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *op);
int bbjd_op_matvec_B(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *op);
int bbjd_op_precond (bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *op);

/* Two preconditioner applies in one pass over the factor (bbjd_precond2_fn):
 * y0 = M^-1 x0, y1 = M^-1 x1. */
int bbjd_op_precond2(bbjd_int n, const bbjd_zomplex *x0, bbjd_zomplex *y0, const bbjd_zomplex *x1,
                     bbjd_zomplex *y1, void *op);

/* A x and B x in one pass over x (bbjd_matvec2_fn): the band rows, the
 * diagonal of B and the dense rows' partial dots in one kernel. */
int bbjd_op_matvec_AB(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *ya, bbjd_zomplex *yb, void *op);

/* bbjd_op_matvec_AB plus r = alpha ya + beta yb and ||r|| -> *rnorm (device)
 * in the same pass and its finish (bbjd_residual_fn). */
int bbjd_op_matvec_ABr(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *ya, bbjd_zomplex *yb, bbjd_zomplex alpha,
                       bbjd_zomplex beta, bbjd_zomplex *r, double *rnorm, void *op);

/* complex64 forms (bbjd_op_c64_fn, bbjd_op2_c64_fn), available when want_c64
 * was set. */
int bbjd_op_matvec_AB_c64(bbjd_int n, const bbjd_cfloat *x, bbjd_cfloat *ya, bbjd_cfloat *yb, void *op);
int bbjd_op_matvec_A_c64(bbjd_int n, const bbjd_cfloat *x, bbjd_cfloat *y, void *op);
int bbjd_op_matvec_B_c64(bbjd_int n, const bbjd_cfloat *x, bbjd_cfloat *y, void *op);
int bbjd_op_precond_c64 (bbjd_int n, const bbjd_cfloat *x, bbjd_cfloat *y, void *op);

/* Block cyclic reduction levels per launch of the exact solve (1..4, default
 * 4; 1: one launch per level).  Process-wide; tuning and tests. */
void bbjd_set_bcr_group(int levels);
/* The exact solve's one-workgroup tail starts at the first reduction level
 * with at most this many rows (default 64); applies to operators created
 * afterwards.  Tuning. */
void bbjd_set_bcr_tail_rows(int rows);
/* 1 (default): the exact solve's last backward reduction launch also forms
 * the dense rows' dots of the Schur border; 0: a separate projection.
 * Tuning and tests. */
void bbjd_set_bcr_border_fuse(int on);
/* 1 (default): for s = 4 and 8 the grouped reduction launches issue every
 * phase's factor loads at launch, each phase on a thread range of its own;
 * 0: one phase after another.  Bit-identical results.  Tuning and tests. */
void bbjd_set_bcr_prefetch(int on);
/* 1 (default): bbjd_op_precond_c64 with the exact preconditioner reads x and
 * writes y in complex64 inside the solve's first reduction launch and its
 * border launches, where the shape allows (dense rows present, s = 4 or 8,
 * at least two levels before the tail); 0: widening and rounding in launches
 * of their own.  Bit-identical results.  Tuning and tests. */
void bbjd_set_c64_direct_io(int on);

#ifdef __cplusplus
}
#endif
#endif /* BBJD_OPS_H */
