/*
 * lapack_host.h -- host LAPACK/BLAS used by the engine (projected QZ, small
 * vectors, preconditioner set-up).  Symbols of scipy's bundled OpenBLAS
 * 0.3.28 (`scipy_` prefix, 32-bit integers) -- the library the reference
 * links (reference setup.py:36-61, include/bbjd_blas.h:30-34).
 */
#ifndef BBJD_LAPACK_HOST_H
#define BBJD_LAPACK_HOST_H
#include <complex.h>

typedef int blas_int;
typedef double complex zcplx;

void scipy_zgemv_(const char *t, const blas_int *m, const blas_int *n, const zcplx *a, const zcplx *A,
                  const blas_int *lda, const zcplx *x, const blas_int *ix, const zcplx *b, zcplx *y,
                  const blas_int *iy);
void scipy_zgges_(const char *jl, const char *jr, const char *srt, void *sel, const blas_int *n, zcplx *A,
                  const blas_int *lda, zcplx *B, const blas_int *ldb, blas_int *sdim, zcplx *al, zcplx *be,
                  zcplx *vsl, const blas_int *ldl, zcplx *vsr, const blas_int *ldr, zcplx *work,
                  const blas_int *lwork, double *rwork, blas_int *bwork, blas_int *info);
void scipy_ztgsen_(const blas_int *ijob, const blas_int *wq, const blas_int *wz, const blas_int *sel,
                   const blas_int *n, zcplx *A, const blas_int *lda, zcplx *B, const blas_int *ldb, zcplx *al,
                   zcplx *be, zcplx *Q, const blas_int *ldq, zcplx *Z, const blas_int *ldz, blas_int *m,
                   double *pl, double *pr, double *dif, zcplx *work, const blas_int *lwork, blas_int *iwork,
                   const blas_int *liwork, blas_int *info);
void scipy_zgetrf_(const blas_int *m, const blas_int *n, zcplx *a, const blas_int *lda, blas_int *ipiv,
                   blas_int *info);
void scipy_zgetri_(const blas_int *n, zcplx *a, const blas_int *lda, const blas_int *ipiv, zcplx *work,
                   const blas_int *lwork, blas_int *info);
void scipy_zggev_(const char *jl, const char *jr, const blas_int *n, zcplx *A, const blas_int *lda, zcplx *B,
                  const blas_int *ldb, zcplx *al, zcplx *be, zcplx *vl, const blas_int *ldvl, zcplx *vr,
                  const blas_int *ldvr, zcplx *work, const blas_int *lwork, double *rwork, blas_int *info);
void scipy_zgetrs_(const char *t, const blas_int *n, const blas_int *nrhs, const zcplx *a, const blas_int *lda,
                   const blas_int *ipiv, zcplx *b, const blas_int *ldb, blas_int *info);
void scipy_zgehrd_(const blas_int *n, const blas_int *ilo, const blas_int *ihi, zcplx *a, const blas_int *lda,
                   zcplx *tau, zcplx *work, const blas_int *lwork, blas_int *info);
void scipy_zhseqr_(const char *job, const char *compz, const blas_int *n, const blas_int *ilo, const blas_int *ihi,
                   zcplx *h, const blas_int *ldh, zcplx *w, zcplx *z, const blas_int *ldz, zcplx *work,
                   const blas_int *lwork, blas_int *info);
void scipy_zunghr_(const blas_int *n, const blas_int *ilo, const blas_int *ihi, zcplx *a, const blas_int *lda,
                   const zcplx *tau, zcplx *work, const blas_int *lwork, blas_int *info);
void scipy_ztrsen_(const char *job, const char *compq, const blas_int *sel, const blas_int *n, zcplx *t,
                   const blas_int *ldt, zcplx *q, const blas_int *ldq, zcplx *w, blas_int *m, double *s,
                   double *sep, zcplx *work, const blas_int *lwork, blas_int *info);
void scipy_ztrsm_(const char *side, const char *uplo, const char *transa, const char *diag, const blas_int *m,
                  const blas_int *n, const zcplx *alpha, const zcplx *a, const blas_int *lda, zcplx *b,
                  const blas_int *ldb);
void scipy_zpotrf_(const char *uplo, const blas_int *n, zcplx *a, const blas_int *lda, blas_int *info);
void scipy_ztrtri_(const char *uplo, const char *diag, const blas_int *n, zcplx *a, const blas_int *lda,
                   blas_int *info);
void scipy_openblas_set_num_threads(int n);
#endif
