/*
 * bbjd_sparse.h -- device operators for general sparse pencils.
 *
 * The reference applies A and B as scipy CSC matvecs and M^-1 through a
 * sparse LU (SuperLU's splu in its tests, FPLUS in its applications;
 * reference bbjd.py:183-219, tests/test_bbjd.py:59-73, README.md:39-44), all
 * on the host.  These are the same three operators as DEVICE-space
 * callbacks for any sparsity: A and B as CSR matrices in HBM, and M^-1 from
 * the factors of a sparse LU computed on the host (once per shift; SuperLU
 * through scipy in bbjd_amd.sparse) whose triangular solves run on the GPU
 * over level sets.  Pass bbjd_sparse_matvec_A / _B / _precond as
 * params.matvec_A / matvec_B / precond and the bbjd_sparse_op* as
 * params.op_data, with callback_space BBJD_CALLBACK_DEVICE.
 */
#ifndef BBJD_SPARSE_H
#define BBJD_SPARSE_H

#include "bbjd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A host CSR matrix: rows n, row offsets ptr[0..n] (ptr[0] = 0), column
 * indices col[ptr[i] .. ptr[i+1]) in [0, ncols) and values val alongside. */
typedef struct {
    bbjd_int            n, ncols;
    const int64_t      *ptr;
    const int32_t      *col;
    const bbjd_zomplex *val;
} bbjd_csr;

/* Factors of a sparse LU of M (SuperLU's convention): Pr M Pc = L U with L
 * lower and U upper triangular (CSR, every row holding its nonzero diagonal
 * entry), (Pr b)[perm_r[i]] = b[i] and (Pc y)[i] = y[perm_c[i]], so that
 * M^-1 b = Pc U^-1 L^-1 Pr b. */
typedef struct {
    bbjd_csr       L, U;
    const int32_t *perm_r, *perm_c;
} bbjd_sparse_lu;

/* Dense rows pinned out of the sparse LU, as FPLUS (the reference's
 * recommended preconditioner, README.md:39-44) does: a row with far more
 * nonzeros than the rest makes a pivoting sparse LU fill densely (SuperLU on
 * a banded pencil with 4 dense rows: nnz(U) 32 M at n = 8000 against 0.12 M
 * for the band part).  With R the d pinned indices (those rows and the same
 * columns) and R' the rest, M11 = M[R', R'] is factored on its own and
 *     y_R  = S^-1 (b_R - M21 M11^-1 b_R'),   S = M22 - M21 Y,
 *     y_R' = M11^-1 b_R' - Y y_R,            Y = M11^-1 M12,
 * with M21 = M[R, R'], M12 = M[R', R], M22 = M[R, R]. */
typedef struct {
    bbjd_int            d;       /* pinned rows, 1 <= d <= 64                */
    const int32_t      *rows;    /* R, increasing, in [0, n)                 */
    const bbjd_zomplex *M21;     /* d x (n - d), row-major                   */
    const bbjd_zomplex *Y;       /* (n - d) x d, column-major                */
    const bbjd_zomplex *Sinv;    /* d x d, column-major                      */
} bbjd_sparse_border;

typedef struct bbjd_sparse_op bbjd_sparse_op;

/* Uploads A, B (NULL: the identity) and the LU factors (NULL: no
 * preconditioner, precond copies) and plans the triangular solves.  stream:
 * hipStream_t for set-up and by default for applies (NULL: the null
 * stream).  Returns NULL on invalid input (shapes, a factor that is not
 * triangular or lacks a nonzero diagonal, a permutation out of range) or a
 * device error. */
bbjd_sparse_op *bbjd_sparse_op_create(const bbjd_csr *A, const bbjd_csr *B, const bbjd_sparse_lu *lu,
                                      void *stream);
/* The same with the preconditioner's dense rows pinned: lu factors M11 (of
 * size n - d, in the order of R', the indices not in border->rows). */
bbjd_sparse_op *bbjd_sparse_op_create_bordered(const bbjd_csr *A, const bbjd_csr *B, const bbjd_sparse_lu *lu,
                                               const bbjd_sparse_border *border, void *stream);
void     bbjd_sparse_op_free(bbjd_sparse_op *op);
size_t   bbjd_sparse_op_bytes(const bbjd_sparse_op *op);   /* device bytes held */
/* level sets of the two triangular solves and the launches each takes */
int      bbjd_sparse_op_levels(const bbjd_sparse_op *op, int64_t *levels_L, int64_t *levels_U,
                               int64_t *launches);

/* DEVICE-space callbacks (bbjd_matvec_fn / bbjd_precond_fn); data is the
 * bbjd_sparse_op*.  They run on bbjd_callback_stream() when called by the
 * engine, else on the op's stream. */
int bbjd_sparse_matvec_A(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *op);
int bbjd_sparse_matvec_B(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *op);
int bbjd_sparse_precond (bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *op);

/* Host logic of the solve plan (no device): the level of every row of a
 * triangular CSR factor (lower = 1: row i depends on columns j < i; lower =
 * 0: on j > i), level[i] = 1 + the largest level it depends on, 0 for none.
 * Returns the number of levels, or -1 when an entry lies on the wrong side
 * of the diagonal. */
int64_t bbjd_sparse_levels(bbjd_int n, const int64_t *ptr, const int32_t *col, int lower, int32_t *level);

#ifdef __cplusplus
}
#endif
#endif /* BBJD_SPARSE_H */
