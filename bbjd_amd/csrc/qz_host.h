/* qz_host.h -- projected QZ step on the host (qz_host.c). */
#ifndef BBJD_QZ_HOST_H
#define BBJD_QZ_HOST_H
#include "../../include/bbjd.h"
#include "lapack_host.h"

typedef struct {
    bbjd_int  maxsub;
    blas_int  lwork, liwork;
    zcplx    *work;
    blas_int *iwork;
    double   *rwork;
    blas_int *bwork, *sel;
    /* cached ZGGES output for unchanged (MA, MB) */
    int       cache_valid;
    bbjd_int  cache_k;
    zcplx    *cSA, *cSB, *cZA, *cZB, *cal, *cbe;
    long      gges_calls, cache_hits;
    /* eigenvalue-only step: its own cache, LU workspace */
    int       ecache_valid;
    bbjd_int  ecache_k;
    zcplx    *eal, *ebe, *lu, *rhs, *tri;
    double   *ebnd;             /* error bound of each cached theta (0: ZGGEV) */
    blas_int *ipiv;
    long      ggev_calls, std_calls, harm_schur_calls;
} bbjd_qz;

int  bbjd_qz_init(bbjd_qz *q, bbjd_int maxsub);
void bbjd_qz_free(bbjd_qz *q);
static inline void bbjd_qz_invalidate(bbjd_qz *q) { q->cache_valid = 0; q->ecache_valid = 0; }
int  bbjd_qz_solve(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB, zcplx *SA, zcplx *SB,
                   zcplx *ZA, zcplx *ZB, zcplx *al, zcplx *be, zcplx sigma, bbjd_target tgt,
                   bbjd_int nkeep, const zcplx *conv, bbjd_int nconv, double tol, bbjd_int *tpos);
int  bbjd_qz_target(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB, zcplx nu0, zcplx mu0, zcplx *al,
                    zcplx *be, zcplx sigma, bbjd_target tgt, const zcplx *conv, bbjd_int nconv, double tol,
                    bbjd_int *tpos);
int  bbjd_qz_schur_harmonic(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB, zcplx nu0, zcplx mu0,
                            zcplx *ZB, zcplx *al, zcplx *be, zcplx sigma, bbjd_target tgt, bbjd_int nkeep,
                            const zcplx *conv, bbjd_int nconv, double tol, bbjd_int *tpos, zcplx *x);
int  bbjd_ritz_vector(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB, zcplx ta, zcplx tb, zcplx *x);
void bbjd_ritz_coefficients(bbjd_int k, bbjd_int ld, bbjd_int p, zcplx ta, zcplx tb, const zcplx *SA,
                            const zcplx *SB, const zcplx *ZB, zcplx *y, zcplx *x);
#endif
