/*
 * orc_jdqz.c -- ORACLE: CPU restatement of the reference JDQZ solver.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker
 * or the timed CPU baseline; the product (libbbjd.so) never calls it.
 *
 * It restates kyle-augustson/BBJD src/jd_core.c step by step with the same
 * host BLAS/LAPACK calls in the same order (scipy OpenBLAS 0.3.28), so on
 * the same inputs it reproduces the reference trajectory to the last bit
 * (pinned against the compiled reference, oracle/_ref, by the tests).
 * Each function names the reference lines it follows.
 *
 * Exported (prefix orc_, no clash with the product):
 *   orc_solve(params, result, trace, trace_cap, n_trace)
 *   orc_block_orthogonalize(n, k, Q, v, h)
 *   orc_project_out(n, k, Q, v)
 *   orc_correction_gmres(...)
 *   orc_zlartg(f, g, &cs, &sn, &r)
 *   orc_result_free(result)
 */
#include "../include/bbjd.h"
#include "orc_lapack.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const lint I1 = 1;

/* One row of the per-iteration trace (the fields the reference prints at
 * verbose 2, src/jd_core.c:894-902, plus the Ritz value parts). */
typedef struct {
    bbjd_int iter, kdim, locked;
    double   lam_re, lam_im, rel_res;
} orc_trace_row;

/* ---- CGS2 with the DGKS test (src/jd_core.c:207-241) ------------------ */
double orc_block_orthogonalize(bbjd_int n, bbjd_int k, const zc *Q, zc *v, zc *h)
{
    lint bn = (lint)n, bk = (lint)k;
    if (k == 0) return scipy_dznrm2_(&bn, v, &I1);
    const zc one = 1.0, mone = -1.0, zero = 0.0;
    double before = scipy_dznrm2_(&bn, v, &I1);
    scipy_zgemv_("C", &bn, &bk, &one, Q, &bn, v, &I1, &zero, h, &I1);
    scipy_zgemv_("N", &bn, &bk, &mone, Q, &bn, h, &I1, &one, v, &I1);
    double after = scipy_dznrm2_(&bn, v, &I1);
    if (after < 0.7071 * before) {          /* second pass, :232-238 */
        scipy_zgemv_("C", &bn, &bk, &one, Q, &bn, v, &I1, &zero, h, &I1);
        scipy_zgemv_("N", &bn, &bk, &mone, Q, &bn, h, &I1, &one, v, &I1);
        after = scipy_dznrm2_(&bn, v, &I1);
    }
    return after;
}

/* ---- v -= Q Q^H v twice, no norm test (src/jd_core.c:269-309) -------- */
void orc_project_out(bbjd_int n, bbjd_int k, const zc *Q, zc *v)
{
    if (k == 0) return;
    lint bn = (lint)n, bk = (lint)k;
    const zc one = 1.0, mone = -1.0, zero = 0.0;
    zc *h = (zc *)malloc((size_t)k * sizeof(zc));
    for (int pass = 0; pass < 2; pass++) {
        scipy_zgemv_("C", &bn, &bk, &one, Q, &bn, v, &I1, &zero, h, &I1);
        scipy_zgemv_("N", &bn, &bk, &mone, Q, &bn, h, &I1, &one, v, &I1);
    }
    free(h);
}

/* ---- Givens with real cs >= 0 (src/jd_core.c:477-511) ---------------- */
void orc_zlartg(zc f, zc g, double *cs, zc *sn, zc *r)
{
    double af = cabs(f), ag = cabs(g);
    if (ag < 1e-300) { *cs = 1.0; *sn = 0.0; *r = f; return; }
    if (af < 1e-300) { *cs = 0.0; *sn = conj(g) / ag; *r = ag; return; }
    double s = af + ag, fs = af / s, gs = ag / s;
    double nrm = s * sqrt(fs * fs + gs * gs);
    *cs = af / nrm;
    *sn = (f / af) * conj(g) / nrm;
    *r  = (f / af) * nrm;
}

/* Selection key of a target (src/jd_core.c:387-396, 445-454). */
static double target_key(bbjd_target t, zc lam, zc sigma)
{
    switch (t) {
    case BBJD_TARGET_LARGEST_MAG:   return -cabs(lam);
    case BBJD_TARGET_SMALLEST_MAG:  return  cabs(lam);
    case BBJD_TARGET_LARGEST_REAL:  return -creal(lam);
    case BBJD_TARGET_SMALLEST_REAL: return  creal(lam);
    case BBJD_TARGET_LARGEST_IMAG:  return -cimag(lam);
    case BBJD_TARGET_SMALLEST_IMAG: return  cimag(lam);
    default:                        return  cabs(lam - sigma);
    }
}

static int matches_converged(zc lam, const zc *conv, bbjd_int nconv, double tol)
{
    for (bbjd_int c = 0; c < nconv; c++)
        if (cabs(lam - conv[c]) < tol * 100) return 1;
    return 0;
}

/* ---- projected QZ + reorder + target (src/jd_core.c:319-462) --------- */
static int orc_projected(bbjd_int k, bbjd_int ld, const zc *MA, const zc *MB,
                         zc *SA, zc *SB, zc *ZA, zc *ZB, zc *al, zc *be,
                         zc sigma, bbjd_target tgt, bbjd_int nkeep,
                         const zc *conv, bbjd_int nconv, double tol, bbjd_int *tpos)
{
    lint bk = (lint)k, lda = (lint)ld, info = 0, sdim = 0;
    for (bbjd_int j = 0; j < k; j++) {
        memcpy(SA + j * ld, MA + j * ld, (size_t)k * sizeof(zc));
        memcpy(SB + j * ld, MB + j * ld, (size_t)k * sizeof(zc));
    }
    double *rwork = malloc((size_t)8 * k * sizeof(double));
    lint *bwork = malloc((size_t)k * sizeof(lint));
    zc wq;
    lint lwork = -1;
    scipy_zgges_("V", "V", "N", NULL, &bk, SA, &lda, SB, &lda, &sdim, al, be,
                 ZA, &lda, ZB, &lda, &wq, &lwork, rwork, bwork, &info);
    if (info) { free(rwork); free(bwork); return -1; }
    lwork = (lint)creal(wq) + 1;
    zc *work = malloc((size_t)lwork * sizeof(zc));
    scipy_zgges_("V", "V", "N", NULL, &bk, SA, &lda, SB, &lda, &sdim, al, be,
                 ZA, &lda, ZB, &lda, work, &lwork, rwork, bwork, &info);
    free(work); free(rwork); free(bwork);
    if (info) return -1;

    lint *sel = calloc((size_t)k, sizeof(lint));
    for (bbjd_int i = 0; i < k; i++) {                       /* :365-374 */
        if (cabs(be[i]) < 1e-300) continue;
        if (matches_converged(al[i] / be[i], conv, nconv, tol)) sel[i] = 1;
    }
    for (bbjd_int pick = 0; pick < nkeep; pick++) {          /* :379-401 */
        bbjd_int best = -1;
        double bv = DBL_MAX;
        for (bbjd_int i = 0; i < k; i++) {
            if (sel[i] || cabs(be[i]) < 1e-300) continue;
            double v = target_key(tgt, al[i] / be[i], sigma);
            if (v < bv) { bv = v; best = i; }
        }
        if (best < 0) break;
        sel[best] = 1;
    }
    lint nsel = 0;
    for (bbjd_int i = 0; i < k; i++) nsel += sel[i];
    if (nsel > 0) {                                          /* :407-425 */
        lint ijob = 0, wantq = 1, wantz = 1, mout = 0, tlw = -1, tli = -1, tiq = 0;
        double pl, pr, dif[2];
        zc twq;
        scipy_ztgsen_(&ijob, &wantq, &wantz, sel, &bk, SA, &lda, SB, &lda, al, be,
                      ZA, &lda, ZB, &lda, &mout, &pl, &pr, dif, &twq, &tlw, &tiq, &tli, &info);
        tlw = (lint)creal(twq) + 1;
        tli = tiq + 1;
        zc *tw = malloc((size_t)tlw * sizeof(zc));
        lint *ti = malloc((size_t)tli * sizeof(lint));
        scipy_ztgsen_(&ijob, &wantq, &wantz, sel, &bk, SA, &lda, SB, &lda, al, be,
                      ZA, &lda, ZB, &lda, &mout, &pl, &pr, dif, tw, &tlw, ti, &tli, &info);
        free(tw); free(ti);
        if (info) { free(sel); return -1; }
    }
    bbjd_int p = 0;                                          /* :430-458 */
    double bv = DBL_MAX;
    for (bbjd_int i = 0; i < k; i++) {
        if (cabs(be[i]) < 1e-300) continue;
        zc lam = al[i] / be[i];
        if (matches_converged(lam, conv, nconv, tol)) continue;
        double v = target_key(tgt, lam, sigma);
        if (v < bv) { bv = v; p = i; }
    }
    *tpos = p;
    free(sel);
    return 0;
}

/* ---- oracle workspace (src/jd_core.c:87-161 sizes) ------------------- */
typedef struct {
    bbjd_int n, maxsub, nev, ims, nlock, kdim;
    zc *V, *W, *AV, *BV, *swap, *MA, *MB, *SA, *SB, *ZA, *ZB;
    zc *gV, *gH, *gcs, *gsn, *gg, *t1, *t2, *t3, *t4, *QR, *QL, *h;
} orc_ws;

static void orc_ws_free(orc_ws *w)
{
    zc **f[] = {&w->V, &w->W, &w->AV, &w->BV, &w->swap, &w->MA, &w->MB, &w->SA, &w->SB,
                &w->ZA, &w->ZB, &w->gV, &w->gH, &w->gcs, &w->gsn, &w->gg, &w->t1, &w->t2,
                &w->t3, &w->t4, &w->QR, &w->QL, &w->h};
    for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); i++) { free(*f[i]); *f[i] = NULL; }
}

static int orc_ws_alloc(orc_ws *w, const bbjd_params *p)
{
    memset(w, 0, sizeof(*w));
    bbjd_int n = p->n, nev = p->nev;
    bbjd_int ms = 3 * nev + 10;
    if (ms < 40) ms = 40;
    w->maxsub = p->maxsub > 0 ? p->maxsub : (ms < n ? ms : n);
    w->ims = p->inner_maxiter > 0 ? p->inner_maxiter : 1;
    w->n = n; w->nev = nev;
    size_t ns = (size_t)n * w->maxsub, ss = (size_t)w->maxsub * w->maxsub;
    w->V = calloc(ns, sizeof(zc)); w->W = calloc(ns, sizeof(zc));
    w->AV = calloc(ns, sizeof(zc)); w->BV = calloc(ns, sizeof(zc));
    w->swap = calloc(ns, sizeof(zc));
    w->MA = calloc(ss, sizeof(zc)); w->MB = calloc(ss, sizeof(zc));
    w->SA = calloc(ss, sizeof(zc)); w->SB = calloc(ss, sizeof(zc));
    w->ZA = calloc(ss, sizeof(zc)); w->ZB = calloc(ss, sizeof(zc));
    w->gV = calloc((size_t)n * (w->ims + 1), sizeof(zc));
    w->gH = calloc((size_t)(w->ims + 1) * w->ims, sizeof(zc));
    w->gcs = calloc((size_t)w->ims, sizeof(zc)); w->gsn = calloc((size_t)w->ims, sizeof(zc));
    w->gg = calloc((size_t)w->ims + 1, sizeof(zc));
    w->t1 = calloc((size_t)n, sizeof(zc)); w->t2 = calloc((size_t)n, sizeof(zc));
    w->t3 = calloc((size_t)n, sizeof(zc)); w->t4 = calloc((size_t)n, sizeof(zc));
    w->QR = calloc((size_t)n * (nev + 1), sizeof(zc));
    w->QL = calloc((size_t)n * (nev + 1), sizeof(zc));
    w->h = calloc((size_t)(2 * w->maxsub + nev + 4), sizeof(zc));
    if (!w->V || !w->W || !w->AV || !w->BV || !w->swap || !w->MA || !w->MB || !w->SA ||
        !w->SB || !w->ZA || !w->ZB || !w->gV || !w->gH || !w->gcs || !w->gsn || !w->gg ||
        !w->t1 || !w->t2 || !w->t3 || !w->t4 || !w->QR || !w->QL || !w->h) {
        orc_ws_free(w);
        return -1;
    }
    return 0;
}

/* ---- deflation column j (src/jd_core.c:516-528) ---------------------- */
static void orc_deflation_append(orc_ws *w, bbjd_int j, const zc *u, const zc *Au,
                                 const zc *Bu, zc nu0, zc mu0)
{
    bbjd_int n = w->n;
    lint bn = (lint)n;
    zc *qr = w->QR + (size_t)j * n, *ql = w->QL + (size_t)j * n;
    scipy_zcopy_(&bn, u, &I1, qr, &I1);
    double nrm = orc_block_orthogonalize(n, j, w->QR, qr, w->h);
    if (nrm > 1e-300) { zc s = 1.0 / nrm; scipy_zscal_(&bn, &s, qr, &I1); }
    for (bbjd_int i = 0; i < n; i++) ql[i] = nu0 * Au[i] + mu0 * Bu[i];
    nrm = orc_block_orthogonalize(n, j, w->QL, ql, w->h);
    if (nrm > 1e-300) { zc s = 1.0 / nrm; scipy_zscal_(&bn, &s, ql, &I1); }
}

/* ---- left-preconditioned GMRES (src/jd_core.c:568-727) --------------- */
int orc_correction_gmres(bbjd_int n, bbjd_int ndefl, const zc *QR, const zc *QL,
                         zc ta, zc tb, const zc *r, zc *z,
                         bbjd_matvec_fn mA, bbjd_matvec_fn mB, bbjd_precond_fn pc,
                         void *od, bbjd_int ims, double tol,
                         zc *V, zc *H, double *cs, zc *sn, zc *g, zc *t1, zc *t2,
                         bbjd_int *steps)
{
    lint bn = (lint)n;
    bbjd_int ldh = ims + 1;
    memset(H, 0, (size_t)(ims + 1) * ims * sizeof(zc));
    for (bbjd_int i = 0; i < n; i++) t1[i] = -r[i];
    orc_project_out(n, ndefl, QL, t1);
    if (pc) { if (pc(n, t1, V, od)) return -1; }
    else scipy_zcopy_(&bn, t1, &I1, V, &I1);
    double beta0 = scipy_dznrm2_(&bn, V, &I1);
    if (beta0 < 1e-300) { memset(z, 0, (size_t)n * sizeof(zc)); return 0; }
    zc ib = 1.0 / beta0;
    scipy_zscal_(&bn, &ib, V, &I1);
    memset(g, 0, (size_t)(ims + 1) * sizeof(zc));
    g[0] = beta0;
    bbjd_int j;
    for (j = 0; j < ims; j++) {
        zc *vj = V + (size_t)j * n, *vj1 = V + (size_t)(j + 1) * n;
        scipy_zcopy_(&bn, vj, &I1, t1, &I1);
        orc_project_out(n, ndefl, QR, t1);
        if (mA(n, t1, t2, od)) return -1;
        scipy_zscal_(&bn, &tb, t2, &I1);
        if (mB(n, t1, vj1, od)) return -1;
        zc na = -ta;
        scipy_zaxpy_(&bn, &na, vj1, &I1, t2, &I1);
        orc_project_out(n, ndefl, QL, t2);
        if (pc) { if (pc(n, t2, vj1, od)) return -1; }
        else scipy_zcopy_(&bn, t2, &I1, vj1, &I1);
        /* Arnoldi by MGS with the DGKS test (:649-669) */
        double wb = scipy_dznrm2_(&bn, vj1, &I1);
        for (bbjd_int i = 0; i <= j; i++) {
            zc hh = orc_zdotc(bn, V + (size_t)i * n, vj1);
            H[i + j * ldh] = hh;
            zc nh = -hh;
            scipy_zaxpy_(&bn, &nh, V + (size_t)i * n, &I1, vj1, &I1);
        }
        double wa = scipy_dznrm2_(&bn, vj1, &I1);
        if (wa < 0.7071 * wb) {
            for (bbjd_int i = 0; i <= j; i++) {
                zc hh = orc_zdotc(bn, V + (size_t)i * n, vj1);
                H[i + j * ldh] += hh;
                zc nh = -hh;
                scipy_zaxpy_(&bn, &nh, V + (size_t)i * n, &I1, vj1, &I1);
            }
            wa = scipy_dznrm2_(&bn, vj1, &I1);
        }
        H[(j + 1) + j * ldh] = wa;
        if (wa < 1e-14 * wb) { j++; break; }                /* :674 */
        zc ih = 1.0 / wa;
        scipy_zscal_(&bn, &ih, vj1, &I1);
        for (bbjd_int i = 0; i < j; i++) {                 /* :680-685 */
            zc h1 = H[i + j * ldh], h2 = H[i + 1 + j * ldh];
            H[i + j * ldh] = cs[i] * h1 + sn[i] * h2;
            H[i + 1 + j * ldh] = -conj(sn[i]) * h1 + cs[i] * h2;
        }
        double c; zc s, rr;
        orc_zlartg(H[j + j * ldh], H[j + 1 + j * ldh], &c, &s, &rr);
        cs[j] = c; sn[j] = s;
        H[j + j * ldh] = rr;
        H[j + 1 + j * ldh] = 0.0;
        zc g1 = g[j], g2 = g[j + 1];
        g[j] = cs[j] * g1 + sn[j] * g2;
        g[j + 1] = -conj(sn[j]) * g1 + cs[j] * g2;
        if (cabs(g[j + 1]) < tol * beta0) { j++; break; }
    }
    bbjd_int m = j;
    if (steps) *steps += m;
    zc *y = g;                                              /* :709-716 */
    for (bbjd_int i = m - 1; i >= 0; i--) {
        for (bbjd_int k2 = i + 1; k2 < m; k2++) y[i] -= H[i + k2 * ldh] * y[k2];
        if (cabs(H[i + i * ldh]) > 1e-300) y[i] /= H[i + i * ldh];
    }
    memset(z, 0, (size_t)n * sizeof(zc));
    for (bbjd_int i = 0; i < m; i++) scipy_zaxpy_(&bn, &y[i], V + (size_t)i * n, &I1, z, &I1);
    orc_project_out(n, ndefl, QR, z);
    return 0;
}

/* ---- the outer iteration (src/jd_core.c:743-1120) -------------------- */
bbjd_status orc_solve(const bbjd_params *P, bbjd_result *R, orc_trace_row *trace,
                      bbjd_int trace_cap, bbjd_int *n_trace, bbjd_int *inner_steps)
{
    orc_ws w;
    bbjd_int n = P->n, nev = P->nev;
    if (n_trace) *n_trace = 0;
    if (inner_steps) *inner_steps = 0;
    memset(R, 0, sizeof(*R));
    if (orc_ws_alloc(&w, P)) { R->status = BBJD_ERR_ALLOC; return BBJD_ERR_ALLOC; }
    bbjd_int ms = w.maxsub;
    bbjd_int minsub = P->minsub > 0 ? P->minsub : (2 * nev + 5 < 20 ? 20 : 2 * nev + 5);
    zc sigma = P->sigma;
    double tol = P->tol;
    lint bn = (lint)n;
    void *od = P->op_data;

    R->eigenvalues = calloc((size_t)nev, sizeof(zc));
    R->eigenvectors = calloc((size_t)n * nev, sizeof(zc));
    R->residuals = calloc((size_t)nev, sizeof(double));
    zc *al = malloc((size_t)ms * sizeof(zc)), *be = malloc((size_t)ms * sizeof(zc));
    bbjd_status st = BBJD_SUCCESS;

    zc nu0 = 0.0, mu0 = 1.0;                                 /* :783-789 */
    if (P->target == BBJD_TARGET_NEAREST) {
        double sc = 1.0 / sqrt(1.0 + cabs(sigma) * cabs(sigma));
        nu0 = sc; mu0 = -sigma * sc;
    } else if (P->target == BBJD_TARGET_SMALLEST_MAG) {
        nu0 = 1.0; mu0 = 0.0;
    }
#define COL(M, j) ((M) + (size_t)(j) * n)
    zc *v0 = w.V;                                            /* :791-820 */
    srand(42);
    if (P->v0) memcpy(v0, P->v0, (size_t)n * sizeof(zc));
    else
        for (bbjd_int i = 0; i < n; i++)
            v0[i] = ((double)rand() / RAND_MAX - 0.5) + ((double)rand() / RAND_MAX - 0.5) * I;
    double nrm = scipy_dznrm2_(&bn, v0, &I1);
    if (nrm < 1e-300) { st = BBJD_ERR_BREAKDOWN; goto done; }
    zc inv = 1.0 / nrm;
    scipy_zscal_(&bn, &inv, v0, &I1);
    if (P->matvec_A(n, v0, w.AV, od) || P->matvec_B(n, v0, w.BV, od)) goto cb_err;
    for (bbjd_int i = 0; i < n; i++) w.W[i] = nu0 * w.AV[i] + mu0 * w.BV[i];
    nrm = scipy_dznrm2_(&bn, w.W, &I1);
    inv = 1.0 / nrm;
    scipy_zscal_(&bn, &inv, w.W, &I1);
    w.MA[0] = orc_zdotc(bn, w.W, w.AV);
    w.MB[0] = orc_zdotc(bn, w.W, w.BV);
    w.kdim = 1;

    for (bbjd_int iter = 0; iter < P->maxiter; iter++) {
        bbjd_int k = w.kdim, locked = w.nlock, p = locked;
        bbjd_int nkeep = minsub - locked;
        if (nkeep < 1) nkeep = 1;
        if (orc_projected(k, ms, w.MA, w.MB, w.SA, w.SB, w.ZA, w.ZB, al, be, sigma,
                          P->target, nkeep, R->eigenvalues, R->nconv, tol, &p)) {
            st = BBJD_ERR_LAPACK;
            goto done;
        }
        zc ta = al[p], tb = be[p];
        zc *u = w.t1;
        {                                                    /* :854-872 */
            zc *y = w.h, *x = w.h + ms;
            y[p] = 1.0;
            for (bbjd_int i = p - 1; i >= 0; i--) {
                zc acc = 0.0;
                for (bbjd_int j = i + 1; j <= p; j++)
                    acc += (tb * w.SA[i + j * ms] - ta * w.SB[i + j * ms]) * y[j];
                zc d = tb * w.SA[i + i * ms] - ta * w.SB[i + i * ms];
                y[i] = (cabs(d) > 1e-300) ? -acc / d : 0.0;
            }
            lint bk = (lint)k, bp1 = (lint)(p + 1), ldz = (lint)ms;
            const zc one = 1.0, zero = 0.0;
            scipy_zgemv_("N", &bk, &bp1, &one, w.ZB, &ldz, y, &I1, &zero, x, &I1);
            scipy_zgemv_("N", &bn, &bk, &one, w.V, &bn, x, &I1, &zero, u, &I1);
        }
        nrm = scipy_dznrm2_(&bn, u, &I1);
        if (nrm < 1e-300) { st = BBJD_ERR_BREAKDOWN; goto done; }
        inv = 1.0 / nrm;
        scipy_zscal_(&bn, &inv, u, &I1);

        zc *Au = w.t2, *Bu = w.t3, *r = w.t4;                /* :879-892 */
        if (P->matvec_A(n, u, Au, od) || P->matvec_B(n, u, Bu, od)) goto cb_err;
        for (bbjd_int i = 0; i < n; i++) r[i] = tb * Au[i] - ta * Bu[i];
        double rn = scipy_dznrm2_(&bn, r, &I1);
        double den = cabs(ta) + cabs(tb);
        double rel = rn / (den > 1e-300 ? den : 1.0);
        zc lam = cabs(tb) > 1e-300 ? ta / tb : ta;
        if (trace && n_trace && *n_trace < trace_cap) {
            orc_trace_row *t = &trace[(*n_trace)++];
            t->iter = iter; t->kdim = k; t->locked = locked;
            t->lam_re = creal(lam); t->lam_im = cimag(lam); t->rel_res = rel;
        }
        if (P->verbose >= 2)
            fprintf(stderr, "  iter %3lld  kdim %3lld  locked %2lld  eig (%+.6e, %+.6e)  res %.3e\n",
                    (long long)iter, (long long)k, (long long)locked, creal(lam), cimag(lam), rel);

        if (rel < tol) {                                     /* :905-932 */
            R->eigenvalues[R->nconv] = lam;
            memcpy(R->eigenvectors + (size_t)R->nconv * n, u, (size_t)n * sizeof(zc));
            R->residuals[R->nconv] = rel;
            R->nconv++;
            if (P->verbose >= 1)
                fprintf(stderr, "  *** converged #%lld: (%+.10e, %+.10e)  res %.3e  iter %lld\n",
                        (long long)R->nconv, creal(lam), cimag(lam), rel, (long long)iter);
            if (R->nconv >= nev) { R->iters = iter + 1; goto done; }
            orc_deflation_append(&w, w.nlock, u, Au, Bu, nu0, mu0);
            w.nlock++;
            continue;
        }

        zc *zcorr = w.t2;                                    /* :944-972 */
        if (P->inner_maxiter > 0) {
            orc_deflation_append(&w, locked, u, Au, Bu, nu0, mu0);
            if (orc_correction_gmres(n, w.nlock + 1, w.QR, w.QL, ta, tb, r, zcorr,
                                     P->matvec_A, P->matvec_B, P->precond, od,
                                     P->inner_maxiter, P->inner_tol, w.gV, w.gH,
                                     (double *)w.gcs, w.gsn, w.gg, w.t3, w.t4, inner_steps)) {
                st = BBJD_ERR_CALLBACK;
                goto done;
            }
        } else {
            if (P->precond) { if (P->precond(n, r, zcorr, od)) { st = BBJD_ERR_CALLBACK; goto done; } }
            else scipy_zcopy_(&bn, r, &I1, zcorr, &I1);
            zc *s = w.t4;
            if (P->precond) { if (P->precond(n, u, s, od)) { st = BBJD_ERR_CALLBACK; goto done; } }
            else scipy_zcopy_(&bn, u, &I1, s, &I1);
            zc uz = orc_zdotc(bn, u, zcorr), us = orc_zdotc(bn, u, s);
            if (cabs(us) > 1e-300) {
                zc c = -(uz / us);
                scipy_zaxpy_(&bn, &c, s, &I1, zcorr, &I1);
            }
        }

        if (k >= ms) {                                       /* :975-1030 */
            bbjd_int keep = minsub < k ? minsub : k - 1;
            if (keep < locked + 1) keep = locked + 1;
            lint bk = (lint)k, bkeep = (lint)keep, lda = (lint)ms;
            const zc one = 1.0, zero = 0.0;
            size_t nk = (size_t)n * keep;
            zc *mats[3] = {w.V, w.AV, w.BV};
            for (int m = 0; m < 3; m++) {
                scipy_zgemm_("N", "N", &bn, &bkeep, &bk, &one, mats[m], &bn, w.ZB, &lda, &zero,
                             w.swap, &bn);
                memcpy(mats[m], w.swap, nk * sizeof(zc));
            }
            for (bbjd_int j = 0; j < keep; j++) {
                zc *wj = COL(w.W, j);
                for (bbjd_int i = 0; i < n; i++) wj[i] = nu0 * COL(w.AV, j)[i] + mu0 * COL(w.BV, j)[i];
                zc *hh = malloc((size_t)(j > 0 ? j : 1) * sizeof(zc));
                nrm = orc_block_orthogonalize(n, j, w.W, wj, hh);
                free(hh);
                if (nrm > 1e-300) { inv = 1.0 / nrm; scipy_zscal_(&bn, &inv, wj, &I1); }
            }
            lint bms = (lint)ms;
            scipy_zgemm_("C", "N", &bkeep, &bkeep, &bn, &one, w.W, &bn, w.AV, &bn, &zero, w.MA, &bms);
            scipy_zgemm_("C", "N", &bkeep, &bkeep, &bn, &one, w.W, &bn, w.BV, &bn, &zero, w.MB, &bms);
            w.kdim = keep;
            k = keep;
        }

        zc *vn = COL(w.V, k);                                /* :1032-1048 */
        scipy_zcopy_(&bn, zcorr, &I1, vn, &I1);
        nrm = orc_block_orthogonalize(n, k, w.V, vn, w.h);
        if (nrm < 1e-14) {
            for (bbjd_int i = 0; i < n; i++)
                vn[i] = ((double)rand() / RAND_MAX - 0.5) + ((double)rand() / RAND_MAX - 0.5) * I;
            nrm = orc_block_orthogonalize(n, k, w.V, vn, w.h);
            if (nrm < 1e-14) { R->iters = iter + 1; st = BBJD_ERR_BREAKDOWN; goto done; }
        }
        inv = 1.0 / nrm;
        scipy_zscal_(&bn, &inv, vn, &I1);
        if (P->matvec_A(n, vn, COL(w.AV, k), od) || P->matvec_B(n, vn, COL(w.BV, k), od)) goto cb_err;
        zc *wn = COL(w.W, k);                                /* :1057-1068 */
        for (bbjd_int i = 0; i < n; i++) wn[i] = nu0 * COL(w.AV, k)[i] + mu0 * COL(w.BV, k)[i];
        nrm = orc_block_orthogonalize(n, k, w.W, wn, w.h);
        if (nrm < 1e-14) {
            scipy_zcopy_(&bn, COL(w.AV, k), &I1, wn, &I1);
            nrm = orc_block_orthogonalize(n, k, w.W, wn, w.h);
        }
        if (nrm > 1e-300) { inv = 1.0 / nrm; scipy_zscal_(&bn, &inv, wn, &I1); }
        {                                                    /* :1085-1104 */
            lint bk1 = (lint)(k + 1);
            const zc one = 1.0, zero = 0.0;
            scipy_zgemv_("C", &bn, &bk1, &one, w.W, &bn, COL(w.AV, k), &I1, &zero, w.MA + k * ms, &I1);
            scipy_zgemv_("C", &bn, &bk1, &one, w.W, &bn, COL(w.BV, k), &I1, &zero, w.MB + k * ms, &I1);
            for (bbjd_int i = 0; i < k; i++) {
                w.MA[k + i * ms] = orc_zdotc(bn, wn, COL(w.AV, i));
                w.MB[k + i * ms] = orc_zdotc(bn, wn, COL(w.BV, i));
            }
        }
        w.kdim = k + 1;
        R->iters = iter + 1;
    }
    if (R->nconv < nev) st = BBJD_MAX_ITER;
done:
    R->status = st;
    free(al); free(be);
    orc_ws_free(&w);
    return st;
cb_err:
    R->status = BBJD_ERR_CALLBACK;
    free(al); free(be);
    orc_ws_free(&w);
    return BBJD_ERR_CALLBACK;
#undef COL
}

void orc_result_free(bbjd_result *r)
{
    if (!r) return;
    free(r->eigenvalues); free(r->eigenvectors); free(r->residuals);
    r->eigenvalues = NULL; r->eigenvectors = NULL; r->residuals = NULL;
}

void orc_set_threads(int t) { scipy_openblas_set_num_threads(t); }
