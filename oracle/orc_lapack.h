/*
 * orc_lapack.h -- ORACLE (test infrastructure only): prototypes of the host
 * BLAS/LAPACK routines of scipy's bundled OpenBLAS (symbols carry the
 * `scipy_` prefix, 32-bit integers).  The reference calls the same routines
 * (include/bbjd_blas.h:40-141 there).
 */
#ifndef ORC_LAPACK_H
#define ORC_LAPACK_H
#include <complex.h>

typedef int lint;
typedef double complex zc;

void   scipy_zaxpy_(const lint *n, const zc *a, const zc *x, const lint *ix, zc *y, const lint *iy);
void   scipy_zscal_(const lint *n, const zc *a, zc *x, const lint *ix);
void   scipy_zcopy_(const lint *n, const zc *x, const lint *ix, zc *y, const lint *iy);
void   scipy_cblas_zdotc_sub(const lint n, const void *x, const lint ix, const void *y, const lint iy, void *out);
double scipy_dznrm2_(const lint *n, const zc *x, const lint *ix);
void   scipy_zgemv_(const char *t, const lint *m, const lint *n, const zc *a, const zc *A, const lint *lda,
                    const zc *x, const lint *ix, const zc *b, zc *y, const lint *iy);
void   scipy_zgemm_(const char *ta, const char *tb, const lint *m, const lint *n, const lint *k, const zc *a,
                    const zc *A, const lint *lda, const zc *B, const lint *ldb, const zc *b, zc *C, const lint *ldc);
void   scipy_zgges_(const char *jl, const char *jr, const char *srt, void *sel, const lint *n,
                    zc *A, const lint *lda, zc *B, const lint *ldb, lint *sdim, zc *al, zc *be,
                    zc *vsl, const lint *ldl, zc *vsr, const lint *ldr, zc *work, const lint *lwork,
                    double *rwork, lint *bwork, lint *info);
void   scipy_ztgsen_(const lint *ijob, const lint *wq, const lint *wz, const lint *sel, const lint *n,
                     zc *A, const lint *lda, zc *B, const lint *ldb, zc *al, zc *be, zc *Q, const lint *ldq,
                     zc *Z, const lint *ldz, lint *m, double *pl, double *pr, double *dif, zc *work,
                     const lint *lwork, lint *iwork, const lint *liwork, lint *info);
void   scipy_zgbtrf_(const lint *m, const lint *n, const lint *kl, const lint *ku, zc *ab, const lint *ldab,
                     lint *ipiv, lint *info);
void   scipy_zgbtrs_(const char *t, const lint *n, const lint *kl, const lint *ku, const lint *nrhs,
                     const zc *ab, const lint *ldab, const lint *ipiv, zc *b, const lint *ldb, lint *info);
void   scipy_zgetrf_(const lint *m, const lint *n, zc *a, const lint *lda, lint *ipiv, lint *info);
void   scipy_zgetrs_(const char *t, const lint *n, const lint *nrhs, const zc *a, const lint *lda,
                     const lint *ipiv, zc *b, const lint *ldb, lint *info);
void   scipy_zgetri_(const lint *n, zc *a, const lint *lda, const lint *ipiv, zc *work, const lint *lwork,
                     lint *info);
void   scipy_openblas_set_num_threads(int n);

static inline zc orc_zdotc(lint n, const zc *x, const zc *y)
{
    zc r;
    scipy_cblas_zdotc_sub(n, x, 1, y, 1, &r);
    return r;
}
#endif
