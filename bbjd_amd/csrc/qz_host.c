/*
 * qz_host.c -- the projected generalized Schur step on the host, FP64.
 *
 * Follows reference src/jd_core.c:319-462 (solve_projected): ZGGES of the
 * k x k projected pencil (MA, MB), selection of the eigenvalues matching
 * converged ones (absolute |lambda - lambda_c| < 100 tol) plus the nkeep best
 * by the target key, ZTGSEN(ijob = 0) to move the selection to the leading
 * block, and the target position as the best unconverged value over ALL
 * positions.  Same LAPACK routines, so the Schur data equal the reference's
 * for equal input.
 *
 * Engine-side differences (results unchanged):
 *  - the left Schur vectors are not accumulated (ZGGES JOBVSL = 'N', ZTGSEN
 *    WANTQ = 0): the iteration only uses SA, SB, ZB, alpha and beta, and
 *    LAPACK applies the same transformations to those either way, so they are
 *    bit-identical to the 'V','V' call while the QZ costs 20-40% less
 *    (ws->ZA is left untouched);
 *  - LAPACK workspaces are sized once per workspace, not malloc'd per call;
 *  - the ZGGES output is cached: on a lock iteration the reference leaves
 *    (MA, MB) unchanged (`continue` at src/jd_core.c:931) so the next ZGGES
 *    would recompute the same factors; the cache replays them and only the
 *    selection and ZTGSEN run again;
 *  - the ordered Schur form is needed only where a thick restart follows
 *    (it keeps the leading Schur vectors).  The other iterations use only
 *    the target Ritz pair: bbjd_qz_target computes the eigenvalues alone
 *    (ZGGEV without vectors: the same alpha, beta as ZGGES's before the
 *    reordering) and picks the target by the same rule, and
 *    bbjd_ritz_vector takes the Ritz vector as the null vector of
 *    tb MA - ta MB by inverse iteration -- in exact arithmetic the vector
 *    the Schur back-substitution gives, up to scale;
 *  - and those eigenvalues come from a standard eigenproblem where the
 *    harmonic test space allows: W is the orthonormal factor of
 *    nu0 AV + mu0 BV (CGS2 column by column, or Cholesky QR at a restart),
 *    so C = nu0 MA + mu0 MB = W^H (nu0 AV + mu0 BV) is upper triangular up
 *    to the loss of orthogonality of W, and MA x = theta MB x is
 *    C^-1 MB x = x / (nu0 theta + mu0) (nu0 != 0) or MB^-1 MA x = theta x
 *    (nu0 = 0, C = MB).  Hessenberg QR of that k x k matrix (ZGEHRD +
 *    ZHSEQR) costs about half of the generalized QZ (ZGGEV).  When C is not
 *    triangular to 1e-12 relative (the w <- Av breakdown fallback) or has a
 *    zero diagonal entry, ZGGEV is used.
 */
#include "qz_host.h"
#include "lapack_host.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

int bbjd_qz_init(bbjd_qz *q, bbjd_int maxsub)
{
    memset(q, 0, sizeof(*q));
    q->maxsub = maxsub;
    size_t ss = (size_t)maxsub * maxsub;
    q->cSA = malloc(ss * sizeof(zcplx)); q->cSB = malloc(ss * sizeof(zcplx));
    q->cZA = malloc(ss * sizeof(zcplx)); q->cZB = malloc(ss * sizeof(zcplx));
    q->cal = malloc((size_t)maxsub * sizeof(zcplx)); q->cbe = malloc((size_t)maxsub * sizeof(zcplx));
    q->rwork = malloc((size_t)8 * maxsub * sizeof(double));
    q->bwork = malloc((size_t)maxsub * sizeof(blas_int));
    q->sel = malloc((size_t)maxsub * sizeof(blas_int));
    /* workspace queries at the largest size */
    blas_int bk = (blas_int)maxsub, lda = bk, info = 0, sdim = 0, lw = -1;
    zcplx wq = 0;
    zcplx *A = calloc(ss, sizeof(zcplx)), *B = calloc(ss, sizeof(zcplx));
    scipy_zgges_("N", "V", "N", NULL, &bk, A, &lda, B, &lda, &sdim, q->cal, q->cbe, q->cZA, &lda, q->cZB,
                 &lda, &wq, &lw, q->rwork, q->bwork, &info);
    q->lwork = (blas_int)creal(wq) + 1;
    blas_int ijob = 0, zero = 0, one = 1, m = 0, tli = -1, tiq = 0;
    double pl, pr, dif[2];
    lw = -1;
    for (bbjd_int i = 0; i < maxsub; i++) q->sel[i] = 1;
    scipy_ztgsen_(&ijob, &zero, &one, q->sel, &bk, A, &lda, B, &lda, q->cal, q->cbe, q->cZA, &lda, q->cZB, &lda,
                  &m, &pl, &pr, dif, &wq, &lw, &tiq, &tli, &info);
    blas_int tl = (blas_int)creal(wq) + 1;
    if (tl > q->lwork) q->lwork = tl;
    q->liwork = tiq + 1 > 1 ? tiq + 1 : 1;
    if (q->lwork < 2 * maxsub + 64) q->lwork = 2 * maxsub + 64;
    lw = -1;
    scipy_zggev_("N", "N", &bk, A, &lda, B, &lda, q->cal, q->cbe, NULL, &bk, NULL, &bk, &wq, &lw, q->rwork, &info);
    tl = (blas_int)creal(wq) + 1;
    if (tl > q->lwork) q->lwork = tl;
    {
        blas_int ilo = 1, ihi = bk;
        lw = -1;
        scipy_zgehrd_(&bk, &ilo, &ihi, A, &lda, q->cal, &wq, &lw, &info);
        tl = (blas_int)creal(wq) + 1;
        if (tl > q->lwork) q->lwork = tl;
        lw = -1;
        scipy_zhseqr_("E", "N", &bk, &ilo, &ihi, A, &lda, q->cal, NULL, &one, &wq, &lw, &info);
        tl = (blas_int)creal(wq) + 1;
        if (tl > q->lwork) q->lwork = tl;
        lw = -1;
        scipy_zhseqr_("S", "V", &bk, &ilo, &ihi, A, &lda, q->cal, B, &lda, &wq, &lw, &info);
        tl = (blas_int)creal(wq) + 1;
        if (tl > q->lwork) q->lwork = tl;
        lw = -1;
        scipy_zunghr_(&bk, &ilo, &ihi, A, &lda, q->cal, &wq, &lw, &info);
        tl = (blas_int)creal(wq) + 1;
        if (tl > q->lwork) q->lwork = tl;
    }
    q->work = malloc((size_t)q->lwork * sizeof(zcplx));
    q->eal = malloc((size_t)maxsub * sizeof(zcplx)); q->ebe = malloc((size_t)maxsub * sizeof(zcplx));
    q->ebnd = malloc((size_t)maxsub * sizeof(double));
    q->tri = malloc(ss * sizeof(zcplx));
    q->lu = malloc(ss * sizeof(zcplx)); q->rhs = malloc((size_t)maxsub * sizeof(zcplx));
    q->ipiv = malloc((size_t)maxsub * sizeof(blas_int));
    q->iwork = malloc((size_t)q->liwork * sizeof(blas_int));
    free(A); free(B);
    if (!q->cSA || !q->cSB || !q->cZA || !q->cZB || !q->cal || !q->cbe || !q->rwork || !q->bwork ||
        !q->sel || !q->work || !q->iwork || !q->eal || !q->ebe || !q->ebnd || !q->lu || !q->rhs || !q->ipiv || !q->tri) {
        bbjd_qz_free(q);
        return -1;
    }
    return 0;
}

void bbjd_qz_free(bbjd_qz *q)
{
    free(q->cSA); free(q->cSB); free(q->cZA); free(q->cZB); free(q->cal); free(q->cbe);
    free(q->rwork); free(q->bwork); free(q->sel); free(q->work); free(q->iwork);
    free(q->eal); free(q->ebe); free(q->ebnd); free(q->lu); free(q->rhs); free(q->ipiv); free(q->tri);
    memset(q, 0, sizeof(*q));
}

/* target key (reference src/jd_core.c:387-396) */
static double key_of(bbjd_target t, zcplx lam, zcplx sigma)
{
    switch (t) {
    case BBJD_TARGET_LARGEST_MAG:   return -cabs(lam);
    case BBJD_TARGET_SMALLEST_MAG:  return cabs(lam);
    case BBJD_TARGET_LARGEST_REAL:  return -creal(lam);
    case BBJD_TARGET_SMALLEST_REAL: return creal(lam);
    case BBJD_TARGET_LARGEST_IMAG:  return -cimag(lam);
    case BBJD_TARGET_SMALLEST_IMAG: return cimag(lam);
    default:                        return cabs(lam - sigma);
    }
}

static int is_converged(zcplx lam, const zcplx *conv, bbjd_int nconv, double tol)
{
    for (bbjd_int c = 0; c < nconv; c++)
        if (cabs(lam - conv[c]) < tol * 100) return 1;
    return 0;
}

int bbjd_qz_solve(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB, zcplx *SA, zcplx *SB, zcplx *ZA,
                  zcplx *ZB, zcplx *al, zcplx *be, zcplx sigma, bbjd_target tgt, bbjd_int nkeep,
                  const zcplx *conv, bbjd_int nconv, double tol, bbjd_int *tpos)
{
    const bbjd_int ld = q->maxsub;
    blas_int bk = (blas_int)k, lda = (blas_int)ld, info = 0;
    const size_t colb = (size_t)k * sizeof(zcplx);
    if (q->cache_valid && q->cache_k == k) {
        q->cache_hits++;
        for (bbjd_int j = 0; j < k; j++) {
            memcpy(SA + j * ld, q->cSA + j * ld, colb); memcpy(SB + j * ld, q->cSB + j * ld, colb);
            memcpy(ZB + j * ld, q->cZB + j * ld, colb);
        }
        memcpy(al, q->cal, colb); memcpy(be, q->cbe, colb);
    } else {
        for (bbjd_int j = 0; j < k; j++) {
            memcpy(SA + j * ld, MA + j * ld, colb);
            memcpy(SB + j * ld, MB + j * ld, colb);
        }
        blas_int sdim = 0, lw = q->lwork;
        scipy_zgges_("N", "V", "N", NULL, &bk, SA, &lda, SB, &lda, &sdim, al, be, ZA, &lda, ZB, &lda, q->work,
                     &lw, q->rwork, q->bwork, &info);
        if (info) return -1;
        q->gges_calls++;
        for (bbjd_int j = 0; j < k; j++) {
            memcpy(q->cSA + j * ld, SA + j * ld, colb); memcpy(q->cSB + j * ld, SB + j * ld, colb);
            memcpy(q->cZB + j * ld, ZB + j * ld, colb);
        }
        memcpy(q->cal, al, colb); memcpy(q->cbe, be, colb);
        q->cache_valid = 1;
        q->cache_k = k;
    }

    blas_int *sel = q->sel;
    for (bbjd_int i = 0; i < k; i++) {
        sel[i] = 0;
        if (cabs(be[i]) < 1e-300) continue;
        if (is_converged(al[i] / be[i], conv, nconv, tol)) sel[i] = 1;
    }
    for (bbjd_int pick = 0; pick < nkeep; pick++) {
        bbjd_int best = -1;
        double bv = DBL_MAX;
        for (bbjd_int i = 0; i < k; i++) {
            if (sel[i] || cabs(be[i]) < 1e-300) continue;
            double v = key_of(tgt, al[i] / be[i], sigma);
            if (v < bv) { bv = v; best = i; }
        }
        if (best < 0) break;
        sel[best] = 1;
    }
    blas_int nsel = 0;
    for (bbjd_int i = 0; i < k; i++) nsel += sel[i];
    if (nsel > 0) {
        blas_int ijob = 0, wq = 0, wz = 1, m = 0, lw = q->lwork, li = q->liwork;
        double pl, pr, dif[2];
        scipy_ztgsen_(&ijob, &wq, &wz, sel, &bk, SA, &lda, SB, &lda, al, be, ZA, &lda, ZB, &lda, &m, &pl, &pr,
                      dif, q->work, &lw, q->iwork, &li, &info);
        if (info) return -1;
    }
    bbjd_int p = 0;
    double bv = DBL_MAX;
    for (bbjd_int i = 0; i < k; i++) {
        if (cabs(be[i]) < 1e-300) continue;
        zcplx lam = al[i] / be[i];
        if (is_converged(lam, conv, nconv, tol)) continue;
        double v = key_of(tgt, lam, sigma);
        if (v < bv) { bv = v; p = i; }
    }
    *tpos = p;
    return 0;
}

/* Ritz coefficients (reference src/jd_core.c:854-870): back-substitution
 * (tb SA - ta SB)[0:p,0:p] y = -(tb SA - ta SB)[0:p,p], y[p] = 1, then
 * x = ZB[0:k, 0:p+1] y by ZGEMV.  y, x need maxsub entries each. */
void bbjd_ritz_coefficients(bbjd_int k, bbjd_int ld, bbjd_int p, zcplx ta, zcplx tb, const zcplx *SA,
                            const zcplx *SB, const zcplx *ZB, zcplx *y, zcplx *x)
{
    y[p] = 1.0;
    for (bbjd_int i = p - 1; i >= 0; i--) {
        zcplx acc = 0.0;
        for (bbjd_int j = i + 1; j <= p; j++) acc += (tb * SA[i + j * ld] - ta * SB[i + j * ld]) * y[j];
        zcplx d = tb * SA[i + i * ld] - ta * SB[i + i * ld];
        y[i] = (cabs(d) > 1e-300) ? -acc / d : 0.0;
    }
    blas_int bk = (blas_int)k, bp1 = (blas_int)(p + 1), ldz = (blas_int)ld, one = 1;
    const zcplx a1 = 1.0, b0 = 0.0;
    scipy_zgemv_("N", &bk, &bp1, &a1, ZB, &ldz, y, &one, &b0, x, &one);
}

/* Eigenvalues of the projected pencil through the standard eigenproblem of
 * the harmonic form (see the file comment), with an error bound per value in
 * q->ebnd; returns 1 when C is not triangular enough to use it.
 * alpha / beta = theta. */
/* M = C^-1 MB (ld k, in q->lu) of the harmonic form, or MB^-1 MA when
 * nu0 = 0, and its Frobenius norm; returns 1 when C is not triangular
 * enough (or singular) to use it. */
static int harmonic_matrix(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB, zcplx nu0, zcplx mu0,
                           double *fro)
{
    const bbjd_int ld = q->maxsub;
    const int inv = cabs(nu0) > 0;               /* C = nu0 MA + mu0 MB; else C = MB */
    zcplx *C = q->tri, *M = q->lu;
    double cmax = 0, lmax = 0, dmin = INFINITY;
    for (bbjd_int j = 0; j < k; j++)
        for (bbjd_int i = 0; i < k; i++) {
            const zcplx c = inv ? nu0 * MA[i + j * ld] + mu0 * MB[i + j * ld] : MB[i + j * ld];
            C[i + j * k] = c;
            const double a = cabs(c);
            if (a > cmax) cmax = a;
            if (i > j && a > lmax) lmax = a;
            if (i == j && a < dmin) dmin = a;
        }
    if (!(cmax > 0) || lmax > 1e-12 * cmax || !(dmin > 1e-300)) return 1;
    /* M = C^-1 MB (or MB^-1 MA), the strictly lower part of C taken as 0 */
    for (bbjd_int j = 0; j < k; j++)
        for (bbjd_int i = 0; i < k; i++) M[i + j * k] = inv ? MB[i + j * ld] : MA[i + j * ld];
    blas_int bk = (blas_int)k;
    const zcplx a1 = 1.0;
    scipy_ztrsm_("L", "U", "N", "N", &bk, &bk, &a1, C, &bk, M, &bk);
    double mfro = 0;
    for (bbjd_int i = 0; i < k * k; i++) mfro += creal(M[i]) * creal(M[i]) + cimag(M[i]) * cimag(M[i]);
    mfro = sqrt(mfro);
    if (!isfinite(mfro)) return 1;
    *fro = mfro;
    return 0;
}

static int eigs_harmonic(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB, zcplx nu0, zcplx mu0)
{
    const int inv = cabs(nu0) > 0;
    zcplx *M = q->lu;
    double mfro = 0;
    if (harmonic_matrix(q, k, MA, MB, nu0, mu0, &mfro)) return 1;
    blas_int bk = (blas_int)k, ilo = 1, ihi = bk, one = 1, lw = q->lwork, info = 0;
    scipy_zgehrd_(&bk, &ilo, &ihi, M, &bk, q->rhs, q->work, &lw, &info);
    if (info) return 1;
    scipy_zhseqr_("E", "N", &bk, &ilo, &ihi, M, &bk, q->eal, NULL, &one, q->work, &lw, &info);
    if (info) return 1;
    /* Accuracy of theta from mu.  The Hessenberg QR gives each mu with an
     * absolute error of about eps ||M||; with nu0 != 0, theta = 1 / (nu0 mu)
     * - mu0 / nu0 moves by |d mu| / (|nu0| |mu|^2).  When a Ritz value sits
     * on sigma (C nearly singular: an eigenvalue at the shift), ||M|| is
     * huge and the values far from sigma lose their digits.  The bound goes
     * with each value; the caller decides whether the values are accurate
     * enough for its choice (pick_target). */
    const double dmu = 4.0 * DBL_EPSILON * mfro;
    for (bbjd_int i = 0; i < k; i++) {
        const zcplx mu = q->eal[i];
        if (!isfinite(creal(mu)) || !isfinite(cimag(mu))) return 1;
        const double amu = cabs(mu);
        if (inv) {
            q->ebnd[i] = amu * cabs(nu0) > 1e-300 ? dmu / (cabs(nu0) * amu * amu) : INFINITY;
            /* mu = 1 / (nu0 theta + mu0): theta = (1 - mu0 mu) / (nu0 mu) */
            q->eal[i] = 1.0 - mu0 * mu;
            q->ebe[i] = nu0 * mu;
        } else {
            q->ebnd[i] = dmu;
            q->ebe[i] = 1.0;
        }
    }
    q->std_calls++;
    return 0;
}

/* Eigenvalues of (MA, MB) by ZGGEV('N', 'N') on copies, error bounds 0. */
static int eigs_ggev(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB)
{
    const bbjd_int ld = q->maxsub;
    const size_t colb = (size_t)k * sizeof(zcplx);
    zcplx *A = q->cSA, *B = q->cSB;   /* scratch: the ZGGES cache is invalid here */
    for (bbjd_int j = 0; j < k; j++) {
        memcpy(A + j * ld, MA + j * ld, colb);
        memcpy(B + j * ld, MB + j * ld, colb);
    }
    blas_int bk = (blas_int)k, lda = (blas_int)ld, one = 1, lw = q->lwork, info = 0;
    scipy_zggev_("N", "N", &bk, A, &lda, B, &lda, q->eal, q->ebe, NULL, &one, NULL, &one, q->work, &lw,
                 q->rwork, &info);
    if (info) return -1;
    for (bbjd_int i = 0; i < k; i++) q->ebnd[i] = 0.0;
    q->cache_valid = 0;
    q->ggev_calls++;
    return 0;
}

static double conv_distance(zcplx lam, const zcplx *conv, bbjd_int nconv)
{
    double m = INFINITY;
    for (bbjd_int c = 0; c < nconv; c++) {
        const double d = cabs(lam - conv[c]);
        if (d < m) m = d;
    }
    return m;
}

/* The target among the cached values (the reference's rule: the best
 * unconverged value by key over all positions, src/jd_core.c:430-458).
 * Returns 0, 2 when the choice is ambiguous for exact values (a tie up to
 * rounding, or no unconverged candidate), or 3 when the values' error
 * bounds could change the choice: a value whose converged match (within
 * 100 tol) is uncertain and could compete, a competitor whose key interval
 * reaches the target's, or a target off by more than 10 tol. */
static int pick_target(const bbjd_qz *q, bbjd_int k, zcplx sigma, bbjd_target tgt, const zcplx *conv,
                       bbjd_int nconv, double tol, bbjd_int *tpos)
{
    const zcplx *al = q->eal, *be = q->ebe;
    const double *eb = q->ebnd;
    bbjd_int p = 0;
    double bv = DBL_MAX, bv2 = DBL_MAX;
    int exact = 1;
    for (bbjd_int i = 0; i < k; i++) {
        if (cabs(be[i]) < 1e-300) continue;
        if (eb[i] != 0.0) exact = 0;
        zcplx lam = al[i] / be[i];
        if (is_converged(lam, conv, nconv, tol)) continue;
        double v = key_of(tgt, lam, sigma);
        if (v < bv) { bv2 = bv; bv = v; p = i; }
        else if (v < bv2) bv2 = v;
    }
    *tpos = p;
    /* The reference scans the positions of its reordered Schur form and the
     * first best value wins; this step's order is the Hessenberg QR's.  The
     * two agree unless the choice is a tie up to rounding or there is no
     * unconverged candidate (the reference then takes position 0 of its
     * ordered form): say so, and the caller decides by the reference's
     * sequence for this iteration (whatever the bounds). */
    if (bv == DBL_MAX) return 2;
    if (bv2 < DBL_MAX) {
        const double scale = fmax(fmax(fabs(bv), fabs(bv2)), 1e-14);
        if (bv2 - bv <= 1e-10 * scale) return 2;
    }
    if (!exact) {
        const double lim = 100.0 * tol, ep = eb[p];
        if (!(ep <= 10.0 * tol) || conv_distance(al[p] / be[p], conv, nconv) - ep < lim) return 3;
        for (bbjd_int i = 0; i < k; i++) {
            if (i == p || cabs(be[i]) < 1e-300) continue;
            const zcplx lam = al[i] / be[i];
            if (conv_distance(lam, conv, nconv) + eb[i] < lim) continue;      /* surely converged */
            if (!(key_of(tgt, lam, sigma) - eb[i] > bv + ep)) return 3;
        }
    }
    return 0;
}

/* The ordered Schur form for a thick restart through the harmonic standard
 * problem: M = C^-1 MB = Z T Z^H (ZGEHRD, ZUNGHR, ZHSEQR), the reference's
 * selection (values within 100 tol of a converged one, then the nkeep best
 * by key) moved to the leading positions in their order (ZTRSEN), the
 * target (best unconverged value over all positions) and its Ritz
 * coefficients x = Z y from T.  M x = mu x is MA x = theta MB x, so the
 * leading columns of Z span the right deflating subspace that the
 * reference's ZGGES + ZTGSEN keeps: the restart keeps the same space.  Z
 * goes to ZB (leading dimension q->maxsub), alpha / beta per position to
 * al, be.  Returns 0, 1 when the generalized QZ is to decide (C not
 * triangular, a choice the values' error bounds could change, a tie for the
 * target or no unconverged candidate), -1 on a LAPACK error. */
int bbjd_qz_schur_harmonic(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB, zcplx nu0, zcplx mu0,
                           zcplx *ZB, zcplx *al, zcplx *be, zcplx sigma, bbjd_target tgt, bbjd_int nkeep,
                           const zcplx *conv, bbjd_int nconv, double tol, bbjd_int *tpos, zcplx *x)
{
    const bbjd_int ld = q->maxsub;
    const int inv = cabs(nu0) > 0;
    zcplx *T = q->lu;
    double mfro = 0;
    q->ecache_valid = 0;                         /* eal, ebe, ebnd are scratch here */
    if (harmonic_matrix(q, k, MA, MB, nu0, mu0, &mfro)) return 1;
    blas_int bk = (blas_int)k, ilo = 1, ihi = bk, lw = q->lwork, ldz = (blas_int)ld, info = 0;
    scipy_zgehrd_(&bk, &ilo, &ihi, T, &bk, q->rhs, q->work, &lw, &info);
    if (info) return -1;
    for (bbjd_int j = 0; j < k; j++) memcpy(ZB + j * ld, T + j * k, (size_t)k * sizeof(zcplx));
    scipy_zunghr_(&bk, &ilo, &ihi, ZB, &ldz, q->rhs, q->work, &lw, &info);
    if (info) return -1;
    scipy_zhseqr_("S", "V", &bk, &ilo, &ihi, T, &bk, q->eal, ZB, &ldz, q->work, &lw, &info);
    if (info) return 1;
    const double dmu = 4.0 * DBL_EPSILON * mfro, lim = 100.0 * tol;
    /* theta, its bound and its distance to the converged values per position */
    double *eb = q->ebnd;
    for (bbjd_int i = 0; i < k; i++) {
        const zcplx mu = T[i + i * k];
        if (!isfinite(creal(mu)) || !isfinite(cimag(mu))) return 1;
        const double amu = cabs(mu);
        if (inv) {
            eb[i] = amu * cabs(nu0) > 1e-300 ? dmu / (cabs(nu0) * amu * amu) : INFINITY;
            al[i] = 1.0 - mu0 * mu;
            be[i] = nu0 * mu;
        } else {
            eb[i] = dmu;
            al[i] = mu;
            be[i] = 1.0;
        }
    }
    /* selection, certain for every value within its bound */
    blas_int *sel = q->sel;
    double kmax = -INFINITY, emax = 0;
    for (bbjd_int i = 0; i < k; i++) {
        sel[i] = 0;
        if (cabs(be[i]) < 1e-300) continue;
        const double m = conv_distance(al[i] / be[i], conv, nconv);
        if (m < lim) sel[i] = 1;
        if (fabs(m - lim) <= eb[i]) return 1;
    }
    for (bbjd_int pick = 0; pick < nkeep; pick++) {
        bbjd_int best = -1;
        double bv = INFINITY;
        for (bbjd_int i = 0; i < k; i++) {
            if (sel[i] || cabs(be[i]) < 1e-300) continue;
            const double v = key_of(tgt, al[i] / be[i], sigma);
            if (v < bv) { bv = v; best = i; }
        }
        if (best < 0) break;
        sel[best] = 1;
        if (bv + eb[best] > kmax + emax) { kmax = bv; emax = eb[best]; }
    }
    if (kmax > -INFINITY)
        for (bbjd_int i = 0; i < k; i++)
            if (!sel[i] && cabs(be[i]) >= 1e-300 && !(key_of(tgt, al[i] / be[i], sigma) - eb[i] > kmax + emax))
                return 1;
    blas_int m = 0;
    double sdum = 0, sepdum = 0;
    scipy_ztrsen_("N", "V", sel, &bk, T, &bk, ZB, &ldz, q->eal, &m, &sdum, &sepdum, q->work, &lw, &info);
    if (info) return -1;
    for (bbjd_int i = 0; i < k; i++) {
        const zcplx mu = T[i + i * k];
        const double amu = cabs(mu);
        if (inv) {
            eb[i] = amu * cabs(nu0) > 1e-300 ? dmu / (cabs(nu0) * amu * amu) : INFINITY;
            al[i] = 1.0 - mu0 * mu;
            be[i] = nu0 * mu;
        } else {
            al[i] = mu;
        }
    }
    /* the target by the reference's scan of the ordered form's positions */
    for (bbjd_int i = 0; i < k; i++) q->ebe[i] = be[i], q->eal[i] = al[i];
    const int rc = pick_target(q, k, sigma, tgt, conv, nconv, tol, tpos);
    if (rc) return 1;
    /* Ritz coefficients: (T(0:p, 0:p) - t_pp I) y = -T(0:p, p), y_p = 1 */
    const bbjd_int p = *tpos;
    zcplx *y = q->rhs;
    const zcplx tpp = T[p + p * k];
    y[p] = 1.0;
    for (bbjd_int i = p - 1; i >= 0; i--) {
        zcplx acc = T[i + p * k];
        for (bbjd_int j = i + 1; j < p; j++) acc += T[i + j * k] * y[j];
        const zcplx d = T[i + i * k] - tpp;
        y[i] = cabs(d) > 1e-300 ? -acc / d : 0.0;
    }
    blas_int bp1 = (blas_int)(p + 1), one = 1;
    const zcplx a1 = 1.0, b0 = 0.0;
    scipy_zgemv_("N", &bk, &bp1, &a1, ZB, &ldz, y, &one, &b0, x, &one);
    q->harm_schur_calls++;
    return 0;
}

/* Eigenvalues of the projected pencil and the target position without the
 * Schur form, cached while the pencil is unchanged (lock iterations): the
 * harmonic standard eigenproblem, else -- or when its error bounds could
 * change the choice of target -- ZGGEV('N', 'N') on copies of (MA, MB).  The
 * target is the best unconverged value by key over all positions, as in
 * bbjd_qz_solve.  Returns 0, -1 on a LAPACK error, or 2 when the choice is
 * ambiguous (a tie up to rounding, or no unconverged candidate): then the
 * ordered Schur form decides, as in the reference. */
int bbjd_qz_target(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB, zcplx nu0, zcplx mu0, zcplx *al,
                   zcplx *be, zcplx sigma, bbjd_target tgt, const zcplx *conv, bbjd_int nconv, double tol,
                   bbjd_int *tpos)
{
    const size_t colb = (size_t)k * sizeof(zcplx);
    if (!(q->ecache_valid && q->ecache_k == k)) {
        if (eigs_harmonic(q, k, MA, MB, nu0, mu0) && eigs_ggev(q, k, MA, MB)) return -1;
        q->cache_valid = 0;
        q->ecache_valid = 1;
        q->ecache_k = k;
    }
    int rc = pick_target(q, k, sigma, tgt, conv, nconv, tol, tpos);
    if (rc == 3) {
        /* the harmonic values are not accurate enough for this choice: the
         * generalized QZ's (cached in their place) decide */
        if (eigs_ggev(q, k, MA, MB)) return -1;
        rc = pick_target(q, k, sigma, tgt, conv, nconv, tol, tpos);
    }
    memcpy(al, q->eal, colb);
    memcpy(be, q->ebe, colb);
    return rc;
}

/* Ritz vector of (MA, MB) for the eigenvalue ta / tb: two steps of inverse
 * iteration with M = tb MA - ta MB (LU with partial pivoting; an exactly
 * zero pivot is replaced by eps ||M||), from a fixed start vector; x is
 * normalised to unit 2-norm (its scale is irrelevant: u = V x is
 * normalised on the device). */
int bbjd_ritz_vector(bbjd_qz *q, bbjd_int k, const zcplx *MA, const zcplx *MB, zcplx ta, zcplx tb, zcplx *x)
{
    const bbjd_int ld = q->maxsub;
    zcplx *M = q->lu;
    double mnorm = 0;
    for (bbjd_int j = 0; j < k; j++)
        for (bbjd_int i = 0; i < k; i++) {
            const zcplx v = tb * MA[i + j * ld] - ta * MB[i + j * ld];
            M[i + j * k] = v;
            const double a = cabs(v);
            if (a > mnorm) mnorm = a;
        }
    blas_int bk = (blas_int)k, one = 1, info = 0;
    scipy_zgetrf_(&bk, &bk, M, &bk, q->ipiv, &info);
    if (info < 0) return -1;
    const double tiny = DBL_EPSILON * (mnorm > 0 ? mnorm : 1.0);
    for (bbjd_int i = 0; i < k; i++)
        if (cabs(M[i + i * k]) < tiny) M[i + i * k] = tiny;
    for (bbjd_int i = 0; i < k; i++) x[i] = 1.0 + 0.5 * I * (double)((i * 7919) % 13) / 13.0;
    for (int step = 0; step < 2; step++) {
        scipy_zgetrs_("N", &bk, &one, M, &bk, q->ipiv, x, &bk, &info);
        if (info) return -1;
        double nrm = 0;
        for (bbjd_int i = 0; i < k; i++) nrm += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
        nrm = sqrt(nrm);
        if (!(nrm > 0) || !isfinite(nrm)) return -1;
        for (bbjd_int i = 0; i < k; i++) x[i] /= nrm;
    }
    return 0;
}
