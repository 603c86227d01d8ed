/*
 * orc_pencil.c -- ORACLE: CPU operators for block-banded pencils with dense
 * tau rows (layout of include/bbjd_ops.h).  TEST INFRASTRUCTURE ONLY (see
 * orc_jdqz.c): the checker for the device operators and the callbacks of
 * the timed CPU baseline.
 *
 * The reference takes these operators from the caller (scipy CSC matvec,
 * bbjd.py:183-205; SuperLU/FPLUS LU, tests/test_bbjd.py:59-73).  Here:
 *   matvec_A: dense band blocks + dense tau rows,
 *   matvec_B: diagonal,
 *   precond : exact LU of M = A - sigma B -- LAPACK zgbtrf/zgbtrs on the
 *             banded leading block, the tau rows through the Schur
 *             complement S = M22 - M21 M11^-1 M12 (zgetrf) -- or the
 *             inverses of the bs x bs diagonal blocks of M (block Jacobi).
 */
#include "../include/bbjd.h"
#include "orc_lapack.h"

#include <stdlib.h>
#include <string.h>

typedef struct {
    bbjd_int n, n1, nb;
    int bs, hw, ntau, pc;
    const zc *band, *tau, *bdiag;
    zc sigma;
    lint kl, ku, ldab;
    zc *ab;  lint *ipiv;      /* zgbtrf factors of M11          */
    zc *Y;                    /* M11^-1 M12, n1 x ntau           */
    zc *S;   lint *spiv;      /* zgetrf factors of the complement */
    zc *bj;                   /* block-Jacobi inverses           */
    long n_apply;
} orc_pencil;

/* A[i][j] of the pencil layout. */
static zc a_entry(const orc_pencil *P, bbjd_int i, bbjd_int j)
{
    if (i >= P->n1) return P->tau[(size_t)(i - P->n1) * P->n + j];
    bbjd_int bi = i / P->bs, bj = j / P->bs, d = bj - bi + P->hw;
    if (d < 0 || d > 2 * P->hw) return 0.0;
    size_t blk = ((size_t)bi * (2 * P->hw + 1) + d) * P->bs * P->bs;
    return P->band[blk + (size_t)(j % P->bs) * P->bs + (i % P->bs)];
}

static zc m_entry(const orc_pencil *P, bbjd_int i, bbjd_int j)
{
    zc v = a_entry(P, i, j);
    if (i == j) v -= P->sigma * P->bdiag[i];
    return v;
}

int orc_matvec_A(bbjd_int n, const zc *x, zc *y, void *data)
{
    const orc_pencil *P = (const orc_pencil *)data;
    int bs = P->bs, hw = P->hw, nd = 2 * hw + 1;
    for (bbjd_int bi = 0; bi < P->nb; bi++) {
        for (int r = 0; r < bs; r++) {
            bbjd_int i = bi * bs + r;
            if (i >= P->n1) break;
            zc acc = 0.0;
            for (int d = 0; d < nd; d++) {
                bbjd_int bj = bi + d - hw;
                if (bj < 0 || bj >= P->nb) continue;
                const zc *blk = P->band + ((size_t)bi * nd + d) * bs * bs;
                for (int c = 0; c < bs; c++) acc += blk[(size_t)c * bs + r] * x[bj * bs + c];
            }
            y[i] = acc;
        }
    }
    for (int t = 0; t < P->ntau; t++) {
        const zc *row = P->tau + (size_t)t * n;
        zc acc = 0.0;
        for (bbjd_int j = 0; j < n; j++) acc += row[j] * x[j];
        y[P->n1 + t] = acc;
    }
    return 0;
}

int orc_matvec_B(bbjd_int n, const zc *x, zc *y, void *data)
{
    const orc_pencil *P = (const orc_pencil *)data;
    for (bbjd_int i = 0; i < n; i++) y[i] = P->bdiag[i] * x[i];
    return 0;
}

int orc_precond(bbjd_int n, const zc *x, zc *y, void *data)
{
    orc_pencil *P = (orc_pencil *)data;
    P->n_apply++;
    if (P->pc == 0) { memcpy(y, x, (size_t)n * sizeof(zc)); return 0; }
    if (P->pc == 2) {
        int bs = P->bs;
        for (bbjd_int bi = 0; bi < P->nb; bi++) {
            const zc *D = P->bj + (size_t)bi * bs * bs;
            for (int r = 0; r < bs; r++) {
                zc acc = 0.0;
                for (int c = 0; c < bs; c++) acc += D[(size_t)c * bs + r] * x[bi * bs + c];
                y[bi * bs + r] = acc;
            }
        }
        return 0;
    }
    lint n1 = (lint)P->n1, one = 1, info = 0;
    memcpy(y, x, (size_t)n1 * sizeof(zc));
    scipy_zgbtrs_("N", &n1, &P->kl, &P->ku, &one, P->ab, &P->ldab, P->ipiv, y, &n1, &info);
    if (info) return -1;
    if (P->ntau > 0) {
        lint nt = P->ntau;
        zc *t = y + P->n1;
        for (int a = 0; a < P->ntau; a++) {
            const zc *row = P->tau + (size_t)a * n;
            zc acc = 0.0;
            for (bbjd_int j = 0; j < P->n1; j++) acc += row[j] * y[j];
            t[a] = x[P->n1 + a] - acc;
        }
        scipy_zgetrs_("N", &nt, &one, P->S, &nt, P->spiv, t, &nt, &info);
        if (info) return -1;
        for (int a = 0; a < P->ntau; a++) {
            const zc *ya = P->Y + (size_t)a * P->n1;
            zc ta = t[a];
            for (bbjd_int j = 0; j < P->n1; j++) y[j] -= ya[j] * ta;
        }
    }
    return 0;
}

void orc_pencil_free(orc_pencil *P)
{
    if (!P) return;
    free(P->ab); free(P->ipiv); free(P->Y); free(P->S); free(P->spiv); free(P->bj);
    free(P);
}

/* pc: 0 none, 1 exact LU, 2 block Jacobi.  Arrays are borrowed. */
orc_pencil *orc_pencil_create(bbjd_int n, int bs, int hw, int ntau, const zc *band,
                              const zc *tau, const zc *bdiag, zc sigma, int pc)
{
    orc_pencil *P = calloc(1, sizeof(*P));
    if (!P) return NULL;
    P->n = n; P->bs = bs; P->hw = hw; P->ntau = ntau; P->pc = pc;
    P->band = band; P->tau = tau; P->bdiag = bdiag; P->sigma = sigma;
    P->nb = n / bs; P->n1 = n - ntau;
    if (pc == 2) {
        P->bj = malloc((size_t)P->nb * bs * bs * sizeof(zc));
        lint lb = bs, info = 0, lw = bs * bs;
        lint *piv = malloc((size_t)bs * sizeof(lint));
        zc *wk = malloc((size_t)lw * sizeof(zc));
        for (bbjd_int bi = 0; bi < P->nb; bi++) {
            zc *D = P->bj + (size_t)bi * bs * bs;
            for (int c = 0; c < bs; c++)
                for (int r = 0; r < bs; r++) D[(size_t)c * bs + r] = m_entry(P, bi * bs + r, bi * bs + c);
            scipy_zgetrf_(&lb, &lb, D, &lb, piv, &info);
            if (!info) scipy_zgetri_(&lb, D, &lb, piv, wk, &lw, &info);
            if (info) { free(piv); free(wk); orc_pencil_free(P); return NULL; }
        }
        free(piv); free(wk);
    } else if (pc == 1) {
        bbjd_int n1 = P->n1;
        lint kl = (lint)((hw + 1) * bs - 1);
        if (kl > n1 - 1) kl = (lint)(n1 - 1);
        P->kl = kl; P->ku = kl; P->ldab = 2 * kl + kl + 1;
        P->ab = calloc((size_t)P->ldab * n1, sizeof(zc));
        P->ipiv = malloc((size_t)n1 * sizeof(lint));
        for (bbjd_int j = 0; j < n1; j++) {
            bbjd_int i0 = j - P->ku < 0 ? 0 : j - P->ku, i1 = j + P->kl > n1 - 1 ? n1 - 1 : j + P->kl;
            for (bbjd_int i = i0; i <= i1; i++)
                P->ab[(size_t)(P->kl + P->ku + i - j) + (size_t)j * P->ldab] = m_entry(P, i, j);
        }
        lint ln1 = (lint)n1, info = 0;
        scipy_zgbtrf_(&ln1, &ln1, &P->kl, &P->ku, P->ab, &P->ldab, P->ipiv, &info);
        if (info) { orc_pencil_free(P); return NULL; }
        if (ntau > 0) {
            lint nt = ntau;
            P->Y = calloc((size_t)n1 * ntau, sizeof(zc));
            for (int a = 0; a < ntau; a++)
                for (bbjd_int i = n1 - 1; i >= 0 && i >= n1 - 1 - P->kl - bs; i--)
                    P->Y[(size_t)a * n1 + i] = m_entry(P, i, n1 + a);
            scipy_zgbtrs_("N", &ln1, &P->kl, &P->ku, &nt, P->ab, &P->ldab, P->ipiv, P->Y, &ln1, &info);
            P->S = malloc((size_t)ntau * ntau * sizeof(zc));
            P->spiv = malloc((size_t)ntau * sizeof(lint));
            for (int s = 0; s < ntau; s++)
                for (int a = 0; a < ntau; a++) {
                    zc acc = 0.0;
                    const zc *row = tau + (size_t)a * n;
                    for (bbjd_int j = 0; j < n1; j++) acc += row[j] * P->Y[(size_t)s * n1 + j];
                    P->S[(size_t)s * ntau + a] = m_entry(P, n1 + a, n1 + s) - acc;
                }
            scipy_zgetrf_(&nt, &nt, P->S, &nt, P->spiv, &info);
            if (info) { orc_pencil_free(P); return NULL; }
        }
    }
    return P;
}

long orc_pencil_napply(const orc_pencil *P) { return P->n_apply; }

/* Dense M = A - sigma B (column-major n x n) for tiny cross-checks. */
void orc_pencil_dense(const orc_pencil *P, zc *A, zc *B)
{
    for (bbjd_int j = 0; j < P->n; j++)
        for (bbjd_int i = 0; i < P->n; i++) {
            A[(size_t)j * P->n + i] = a_entry(P, i, j);
            B[(size_t)j * P->n + i] = i == j ? P->bdiag[i] : 0.0;
        }
}
