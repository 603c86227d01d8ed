/*
 * ops_host.c -- device operators for block-banded pencils with dense tau
 * rows (include/bbjd_ops.h): set-up on the host, applies on the GPU.
 *
 * The reference leaves these to the caller (scipy CSC matvec, bbjd.py:
 * 183-205; SuperLU / FPLUS exact LU, tests/test_bbjd.py:59-73, README.md:
 * 39-44).  Here the exact preconditioner of M = A - sigma B is
 *   - block cyclic reduction (BCR) of the leading n1 = n - ntau rows, grouped
 *     into super-blocks of s = hw * bs rows so that they form a block-
 *     tridiagonal matrix (padded to a multiple of s with identity rows), and
 *   - the ntau dense rows through the Schur complement
 *     S = M22 - M21 M11^-1 M12, applied as
 *     y1 = M11^-1 b1, y2 = S^-1 (b2 - M21 y1), x1 = y1 - (M11^-1 M12) y2.
 * BCR eliminates without pivoting between super-blocks (each s x s pivot
 * block is inverted with partial pivoting).  That is exact in exact
 * arithmetic; its rounding is checked against LAPACK's pivoted band LU in
 * the tests.
 */
#include "../../include/bbjd_ops.h"
#include "engine.h"
#include "ops_dev.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zcx;
typedef struct { float re, im; } c64;
#define HIPOK(x) ((x) == hipSuccess)

struct bbjd_op {
    bbjd_int n, n1, nb;
    int bs, hw, ntau, pc, nd, want_c64;
    hipStream_t stream;
    zcx sigma;
    void *dblock;
    size_t dbytes;
    void *band, *tauc, *bdiag;          /* complex128 operator data          */
    void *band_c, *tauc_c, *bdiag_c;    /* complex64 copies                  */
    void *bj, *bj_c;                    /* block-Jacobi inverses             */
    /* BCR */
    int s;
    bbjd_int N;
    int L;
    void *bcr;                          /* factor storage                    */
    bbjd_int *even_off, *odd_off;       /* per level, in s x s blocks        */
    bbjd_int final_off;
    int l0;                             /* first level run by the tail kernel */
    void *doff;                         /* device copy of the level offsets  */
    void *bvec, *xvec;                  /* padded work vectors (N s)         */
    void *bvec2;                        /* second right-hand-side buffer: fused
                                           forward levels read one, write the other */
    void *Y, *Sinv, *small;             /* border                            */
    bbjd_red rd;                        /* tau-row dot scratch               */
    void *zt1, *zt2;                    /* complex128 scratch for c64 applies */
};

/* ---- small dense helpers (column-major s x s) ---------------------- */
static void mm(int s, const zcx *A, const zcx *B, zcx *C)      /* C = A B */
{
    for (int j = 0; j < s; j++) {
        zcx *cj = C + (size_t)j * s;
        for (int i = 0; i < s; i++) cj[i] = 0;
        for (int l = 0; l < s; l++) {
            const zcx b = B[(size_t)j * s + l];
            const zcx *al = A + (size_t)l * s;
            for (int i = 0; i < s; i++) cj[i] += al[i] * b;
        }
    }
}
static void mm_sub(int s, const zcx *A, const zcx *B, zcx *C)  /* C -= A B */
{
    for (int j = 0; j < s; j++) {
        zcx *cj = C + (size_t)j * s;
        for (int l = 0; l < s; l++) {
            const zcx b = B[(size_t)j * s + l];
            const zcx *al = A + (size_t)l * s;
            for (int i = 0; i < s; i++) cj[i] -= al[i] * b;
        }
    }
}
/* in-place inverse by Gauss-Jordan with partial pivoting; -1 if singular */
static int zinv(int s, zcx *A, zcx *work /* s*s */, int *piv)
{
    zcx *E = work;
    for (int j = 0; j < s; j++)
        for (int i = 0; i < s; i++) E[(size_t)j * s + i] = i == j;
    for (int k = 0; k < s; k++) {
        int p = k;
        double best = cabs(A[(size_t)k * s + k]);
        for (int i = k + 1; i < s; i++) {
            double v = cabs(A[(size_t)k * s + i]);
            if (v > best) { best = v; p = i; }
        }
        if (!(best > 0.0) || !isfinite(best)) return -1;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < s; j++) {
                zcx t = A[(size_t)j * s + k]; A[(size_t)j * s + k] = A[(size_t)j * s + p]; A[(size_t)j * s + p] = t;
                t = E[(size_t)j * s + k]; E[(size_t)j * s + k] = E[(size_t)j * s + p]; E[(size_t)j * s + p] = t;
            }
        const zcx inv = 1.0 / A[(size_t)k * s + k];
        for (int j = 0; j < s; j++) { A[(size_t)j * s + k] *= inv; E[(size_t)j * s + k] *= inv; }
        for (int i = 0; i < s; i++) {
            if (i == k) continue;
            const zcx f = A[(size_t)k * s + i];
            if (f == 0) continue;
            for (int j = 0; j < s; j++) {
                A[(size_t)j * s + i] -= f * A[(size_t)j * s + k];
                E[(size_t)j * s + i] -= f * E[(size_t)j * s + k];
            }
        }
    }
    memcpy(A, E, (size_t)s * s * sizeof(zcx));
    return 0;
}

/* A[i][j] of the pencil */
static zcx a_at(const bbjd_pencil_desc *d, bbjd_int i, bbjd_int j)
{
    const bbjd_int n1 = d->n - d->ntau;
    if (i >= n1) return ((const zcx *)d->tau)[(size_t)(i - n1) * d->n + j];
    const bbjd_int bi = i / d->bs, bj = j / d->bs, dd = bj - bi + d->hw;
    if (dd < 0 || dd > 2 * d->hw) return 0;
    return ((const zcx *)d->band)[((size_t)bi * (2 * d->hw + 1) + dd) * d->bs * d->bs + (size_t)(j % d->bs) * d->bs +
                                  (i % d->bs)];
}
static zcx m_at(const bbjd_pencil_desc *d, zcx sigma, bbjd_int i, bbjd_int j)
{
    zcx v = a_at(d, i, j);
    if (i == j) v -= sigma * ((const zcx *)d->bdiag)[i];
    return v;
}

/* super-block (I, J) of M11 padded with identity */
static void m11_block(const bbjd_pencil_desc *d, zcx sigma, int s, bbjd_int n1, bbjd_int bI, bbjd_int bJ, zcx *B)
{
    for (int c = 0; c < s; c++)
        for (int r = 0; r < s; r++) {
            const bbjd_int i = bI * s + r, j = bJ * s + c;
            zcx v = 0;
            if (i < n1 && j < n1) v = m_at(d, sigma, i, j);
            else if (i == j) v = 1.0;
            B[(size_t)c * s + r] = v;
        }
}

/* The one-workgroup tail starts at the first level with at most this many
 * rows (M s); set before an operator is created.  With the launched levels
 * fused four per launch a level there costs little, so the serial tail is
 * kept short: 64 measured best on MI355X (s = 8, one-vector solve: 81 us
 * at n = 2e5 against 86 with 256, level at n = 2e4).  Tunable. */
static int g_bcr_tail_rows = 64;
void bbjd_set_bcr_tail_rows(int rows) { g_bcr_tail_rows = rows < 1 ? 64 : rows; }

/* Host BCR factorisation; fills op->bcr (host image) and the offsets. */
static int bcr_factor(struct bbjd_op *op, const bbjd_pencil_desc *d, zcx **host_out, size_t *nblocks_out)
{
    const int s = op->s;
    const bbjd_int N = op->N;
    const size_t ss = (size_t)s * s;
    int L = 0;
    while (((bbjd_int)1 << L) < N) L++;
    op->L = L;
    op->even_off = calloc((size_t)L + 1, sizeof(bbjd_int));
    op->odd_off = calloc((size_t)L + 1, sizeof(bbjd_int));
    size_t nbk = 0;
    for (int l = 0; l < L; l++) {
        const bbjd_int M = (N + ((bbjd_int)1 << l) - 1) >> l;
        op->even_off[l] = (bbjd_int)nbk;
        nbk += (size_t)2 * ((M + 1) / 2);
        op->odd_off[l] = (bbjd_int)nbk;
        nbk += (size_t)3 * (M / 2);
    }
    op->final_off = (bbjd_int)nbk;
    nbk += 1;
    /* The one-workgroup tail starts at the first level with at most
     * g_bcr_tail_rows rows (M s): a level costs it about 0.17 us per block row
     * (MI355X, s = 8, two right-hand sides), so wider levels run faster in
     * the launched groups. */
    op->l0 = L;
    for (int l = 0; l < L; l++) {
        const bbjd_int M = (N + ((bbjd_int)1 << l) - 1) >> l;
        if (M * s <= g_bcr_tail_rows) { op->l0 = l; break; }
    }
    zcx *H = malloc(nbk * ss * sizeof(zcx));
    zcx *Lc = malloc((size_t)N * ss * sizeof(zcx)), *Dc = malloc((size_t)N * ss * sizeof(zcx)),
        *Uc = malloc((size_t)N * ss * sizeof(zcx));
    if (!H || !Lc || !Dc || !Uc) { free(H); free(Lc); free(Dc); free(Uc); return -1; }
    const bbjd_int n1 = op->n1;
#pragma omp parallel for schedule(static)
    for (bbjd_int bI = 0; bI < N; bI++) {
        m11_block(d, op->sigma, s, n1, bI, bI, Dc + bI * ss);
        if (bI > 0) m11_block(d, op->sigma, s, n1, bI, bI - 1, Lc + bI * ss);
        else memset(Lc, 0, ss * sizeof(zcx));
        if (bI + 1 < N) m11_block(d, op->sigma, s, n1, bI, bI + 1, Uc + bI * ss);
        else memset(Uc + bI * ss, 0, ss * sizeof(zcx));
    }
    int bad = 0;
    for (int l = 0; l < L; l++) {
        const bbjd_int h = (bbjd_int)1 << l, M = (N + h - 1) >> l;
        zcx *odd = H + (size_t)op->odd_off[l] * ss, *even = H + (size_t)op->even_off[l] * ss;
#pragma omp parallel for schedule(static) reduction(| : bad)
        for (bbjd_int j = 0; j < M / 2; j++) {
            const bbjd_int i = (2 * j + 1) * h;
            zcx *Dinv = odd + (size_t)j * 3 * ss, *Pm = Dinv + ss, *Qm = Pm + ss;
            zcx work[64 * 64];
            int piv[64];
            memcpy(Dinv, Dc + i * ss, ss * sizeof(zcx));
            if (zinv(s, Dinv, work, piv)) bad |= 1;
            mm(s, Dinv, Lc + i * ss, Pm);
            mm(s, Dinv, Uc + i * ss, Qm);
        }
#pragma omp parallel for schedule(static)
        for (bbjd_int j = 0; j < (M + 1) / 2; j++) {
            const bbjd_int m = 2 * j, i = m * h;
            zcx *al = even + (size_t)j * 2 * ss, *ga = al + ss;
            zcx nl[64 * 64], nu[64 * 64];
            memset(nl, 0, ss * sizeof(zcx));
            memset(nu, 0, ss * sizeof(zcx));
            if (m >= 1) {
                const zcx *Dl = odd + (size_t)(j - 1) * 3 * ss;   /* odd m-1 -> j-1 */
                mm(s, Lc + i * ss, Dl, al);
                mm_sub(s, al, Uc + (i - h) * ss, Dc + i * ss);
                mm_sub(s, al, Lc + (i - h) * ss, nl);
            } else memset(al, 0, ss * sizeof(zcx));
            if (m + 1 < M) {
                const zcx *Dr = odd + (size_t)j * 3 * ss;         /* odd m+1 -> j */
                mm(s, Uc + i * ss, Dr, ga);
                mm_sub(s, ga, Lc + (i + h) * ss, Dc + i * ss);
                mm_sub(s, ga, Uc + (i + h) * ss, nu);
            } else memset(ga, 0, ss * sizeof(zcx));
            memcpy(Lc + i * ss, nl, ss * sizeof(zcx));
            memcpy(Uc + i * ss, nu, ss * sizeof(zcx));
        }
    }
    {
        zcx *F = H + (size_t)op->final_off * ss;
        zcx work[64 * 64];
        int piv[64];
        memcpy(F, Dc, ss * sizeof(zcx));
        if (zinv(s, F, work, piv)) bad |= 1;
    }
    free(Lc); free(Dc); free(Uc);
    if (bad) { free(H); return -1; }
    *host_out = H;
    *nblocks_out = nbk;
    return 0;
}

static hipStream_t op_stream(const struct bbjd_op *op)
{
    hipStream_t s = (hipStream_t)bbjd_callback_stream();
    return s ? s : op->stream;
}

/* Levels per fused BCR launch (1: one launch per level). */
static int g_bcr_group = 4;
void bbjd_set_bcr_group(int levels) { g_bcr_group = levels < 1 ? 1 : (levels > 4 ? 4 : levels); }
/* 1: the last backward group also forms the dense rows' partial dots for the
 * Schur border (no separate projection and finish) */
static int g_bcr_border_fuse = 1;
void bbjd_set_bcr_border_fuse(int on) { g_bcr_border_fuse = on != 0; }
/* complex64 applies of the exact preconditioner read and write complex64 in
 * the solve's own first and last launches (1) or convert in launches of
 * their own (0) */
static int g_c64_io = 1;
void bbjd_set_c64_direct_io(int on) { g_c64_io = on != 0; }

/* BCR solve of nrhs right-hand sides into op->xvec: the large levels in
 * launches of up to g_bcr_group levels, the levels from the tail level on,
 * the final solve and their backward levels in one workgroup.  With x0 (x1)
 * the right-hand sides are the caller's n1-row vectors, read by the first
 * levels (or copied into op->bvec when every level is in the tail);
 * without, they are already in op->bvec (vector q at q N s).
 *
 * A fused forward group reads one right-hand-side buffer and writes the
 * rows at multiples of 2^l into the other (its halo must see the values
 * from before the group), so the current buffer alternates between bvec
 * and bvec2; the odd rows of level l are final once its group has run and
 * stay in the buffer it wrote (later groups write only multiples of higher
 * powers of two), which the backward level reads (bl[l]). */
/* in_c64: x0 is complex64 (one right-hand side); -3 before any launch when
 * the first reduction launch cannot read it */
static int bcr_solve(struct bbjd_op *op, hipStream_t st, int nrhs, const void *x0, const void *x1, int want_dots,
                     int64_t *dots_nblk, int in_c64)
{
    if (dots_nblk) *dots_nblk = 0;
    const int s = op->s;
    const size_t ss = (size_t)s * s;
    const int64_t vs = (int64_t)op->N * s;
    int rc = 0;
    const int lt = bbjd_o_bcr_tail_level(s, op->N, op->l0, op->L, nrhs);
    if (in_c64 && (!x0 || s > 64 || (lt < g_bcr_group ? lt : g_bcr_group) < 2)) return -3;
    void *buf[2] = {op->bvec, op->bvec2};
    int cur = 0;
    int bl[64];
    const void *lv[4], *lb[4];
    int l = 0;
    if (x0 && lt >= 1 && s <= 64) {
        const int G = lt < g_bcr_group ? lt : g_bcr_group;
        int e = -3;
        if (G >= 2) {
            for (int g = 0; g < G; g++) lv[g] = (const zcx *)op->bcr + (size_t)op->even_off[g] * ss;
            e = bbjd_o_bcr_fwd_multi(st, s, op->N, op->n1, 0, G, lv, x0, x1, NULL, buf[0], nrhs, vs, in_c64);
            if (e == 0) l = G;
            else if (in_c64) return e;
            else if (e != -3) rc |= e;
        }
        if (e == -3) {
            rc |= bbjd_o_bcr_fwd0(st, s, op->N, op->n1, (const zcx *)op->bcr + (size_t)op->even_off[0] * ss, x0,
                                  x1, buf[0], nrhs, vs);
            l = 1;
        }
        for (int k = 0; k < l; k++) bl[k] = 0;
    } else if (x0) {
        rc |= bbjd_copy_async(op->bvec, x0, (size_t)op->n1 * 16, hipMemcpyDeviceToDevice, st) != 0;
        if (nrhs == 2)
            rc |= bbjd_copy_async((zcx *)op->bvec + vs, x1, (size_t)op->n1 * 16, hipMemcpyDeviceToDevice, st) != 0;
    }
    while (l < lt) {
        const int G = lt - l < g_bcr_group ? lt - l : g_bcr_group;
        if (G >= 2) {
            for (int g = 0; g < G; g++) lv[g] = (const zcx *)op->bcr + (size_t)op->even_off[l + g] * ss;
            const int e = bbjd_o_bcr_fwd_multi(st, s, op->N, op->n1, l, G, lv, NULL, NULL, buf[cur], buf[1 - cur],
                                               nrhs, vs, 0);
            if (e == 0) {
                cur = 1 - cur;
                for (int g = 0; g < G; g++) bl[l + g] = cur;
                l += G;
                continue;
            }
            if (e != -3) { rc |= e; break; }
        }
        const bbjd_int h = (bbjd_int)1 << l, M = (op->N + h - 1) >> l;
        rc |= bbjd_o_bcr_fwd(st, s, M, h, (M + 1) / 2, (const zcx *)op->bcr + (size_t)op->even_off[l] * ss,
                             buf[cur], nrhs, vs);
        bl[l] = cur;
        l++;
    }
    rc |= bbjd_o_bcr_tail(st, s, op->N, lt, op->L, (const int64_t *)op->doff, op->bcr, buf[cur], op->xvec, nrhs,
                          vs);
    for (int l = lt - 1; l >= 0;) {
        const int G = l + 1 < g_bcr_group ? l + 1 : g_bcr_group;
        if (G >= 2) {
            const int lo = l - G + 1;
            for (int g = 0; g < G; g++) {
                lv[g] = (const zcx *)op->bcr + (size_t)op->odd_off[lo + g] * ss;
                lb[g] = buf[bl[lo + g]];
            }
            const int dots = want_dots && lo == 0;
            const int e = bbjd_o_bcr_bwd_multi(st, s, op->N, lo, G, lv, lb, op->xvec, nrhs, vs,
                                               dots ? op->tauc : NULL, op->ntau, op->n1, op->n, op->rd.part,
                                               op->rd.part_cap, dots ? dots_nblk : NULL);
            if (e == 0) { l = lo - 1; continue; }
            if (e != -3) { rc |= e; break; }
        }
        const bbjd_int h = (bbjd_int)1 << l, M = (op->N + h - 1) >> l;
        rc |= bbjd_o_bcr_bwd(st, s, M, h, M / 2, (const zcx *)op->bcr + (size_t)op->odd_off[l] * ss, buf[bl[l]],
                             op->xvec, nrhs, vs);
        l--;
    }
    return rc;
}

/* y_q = M^-1 x_q for nrhs (1 or 2) vectors */
/* c64: x0 and y0 are complex64 (one right-hand side), read and written by
 * the first reduction launch and the border kernels; -3 before any launch
 * when this operator's shape does not allow it (no dense rows, s > 8, too
 * few levels before the tail) */
static int exact_apply(struct bbjd_op *op, hipStream_t st, int nrhs, const void *x0, void *y0, const void *x1,
                       void *y1, int c64)
{
    int rc = 0;
    const bbjd_int n1 = op->n1;
    const int64_t vs = (int64_t)op->N * op->s;
    int64_t dnb = 0;
    if (c64 && (nrhs != 1 || op->ntau <= 0)) return -3;
    {
        const int e = bcr_solve(op, st, nrhs, x0, x1, op->ntau > 0 && g_bcr_border_fuse, &dnb, c64);
        if (c64 && e == -3) return -3;
        rc |= e;
    }
    if (op->ntau == 0) {
        rc |= bbjd_copy_async(y0, op->xvec, (size_t)n1 * 16, hipMemcpyDeviceToDevice, st) != 0;
        if (nrhs == 2)
            rc |= bbjd_copy_async(y1, (zcx *)op->xvec + vs, (size_t)n1 * 16, hipMemcpyDeviceToDevice, st) != 0;
        return rc;
    }
    const int nt = op->ntau;
    zcx *sm = (zcx *)op->small;   /* dots 2 nt | y2 2 nt */
    if (!dnb)
        rc |= bbjd_k_gemv_c(st, n1, nt, op->tauc, op->n, op->xvec, nrhs == 2 ? (zcx *)op->xvec + vs : NULL, sm,
                            sm + nt, &op->rd);
    /* the dense rows' part of x and y: complex64 elements are 8 bytes */
    const void *b2 = c64 ? (const void *)((const char *)x0 + (size_t)n1 * 8) : (const void *)((const zcx *)x0 + n1);
    void *yt = c64 ? (void *)((char *)y0 + (size_t)n1 * 8) : (void *)((zcx *)y0 + n1);
    rc |= bbjd_o_border_small(st, nt, b2, nrhs == 2 ? (const zcx *)x1 + n1 : NULL, sm, sm + nt, op->Sinv,
                              sm + 2 * nt, yt, nrhs == 2 ? (zcx *)y1 + n1 : NULL, nrhs, dnb ? op->rd.part : NULL, dnb,
                              c64);
    rc |= bbjd_o_border_update(st, n1, nt, op->xvec, vs, op->Y, sm + 2 * nt, y0, y1, nrhs, c64);
    return rc;
}

/* ---- device layout --------------------------------------------------- */
enum { D_BAND, D_TAUC, D_BDIAG, D_BANDC, D_TAUCC, D_BDIAGC, D_BJ, D_BJC, D_BCR, D_BV, D_XV, D_Y, D_SINV, D_SMALL,
       D_PART, D_PNORM, D_RSC, D_ZT1, D_ZT2, D_OFF, D_BV2, D_END };

static size_t bcr_block_count(bbjd_int N)
{
    int L = 0;
    while (((bbjd_int)1 << L) < N) L++;
    size_t nbk = 0;
    for (int l = 0; l < L; l++) {
        const bbjd_int M = (N + ((bbjd_int)1 << l) - 1) >> l;
        nbk += (size_t)2 * ((M + 1) / 2) + (size_t)3 * (M / 2);
    }
    return nbk + 1;
}

static size_t op_layout(const bbjd_pencil_desc *d, int pc, int want_c64, bbjd_int N, int s, size_t bcr_blocks,
                        size_t *sz, size_t *off)
{
    const bbjd_int n = d->n, n1 = d->n - d->ntau, nb = d->n / d->bs;
    const size_t nbandel = (size_t)nb * (2 * d->hw + 1) * d->bs * d->bs;
    memset(sz, 0, sizeof(size_t) * D_END);
    sz[D_BAND] = nbandel * 16;
    sz[D_TAUC] = (size_t)n * d->ntau * 16;
    sz[D_BDIAG] = (size_t)n * 16;
    if (want_c64) { sz[D_BANDC] = nbandel * 8; sz[D_TAUCC] = (size_t)n * d->ntau * 8; sz[D_BDIAGC] = (size_t)n * 8; }
    if (pc == BBJD_PC_BLOCK_JACOBI) {
        sz[D_BJ] = (size_t)nb * d->bs * d->bs * 16;
        if (want_c64) sz[D_BJC] = (size_t)nb * d->bs * d->bs * 8;
    }
    if (pc == BBJD_PC_EXACT_LU) {
        sz[D_BCR] = bcr_blocks * s * s * 16;
        sz[D_BV] = sz[D_XV] = sz[D_BV2] = (size_t)2 * N * s * 16;      /* two right-hand sides */
        sz[D_OFF] = (size_t)(2 * 64 + 2) * 8;
        sz[D_Y] = (size_t)n1 * d->ntau * 16;
        sz[D_SINV] = (size_t)d->ntau * d->ntau * 16 + 16;
    }
    sz[D_SMALL] = (size_t)(4 * d->ntau + 16) * 16;
    sz[D_PART] = (size_t)bbjd_red_part_entries(n, d->ntau > 0 ? d->ntau + 2 : 2) * 16;
    sz[D_PNORM] = ((size_t)(n + BBJD_TILE_MIN_ROWS - 1) / BBJD_TILE_MIN_ROWS + 64) * 8;
    sz[D_RSC] = 512 + BBJD_CNT_SLOTS * 4 + 1024;
    if (want_c64 && pc != BBJD_PC_NONE) sz[D_ZT1] = sz[D_ZT2] = (size_t)n * 16;
    size_t tot = 0;
    for (int i = 0; i < D_END; i++) { off[i] = tot; tot += (sz[i] + 255) / 256 * 256; }
    return tot;
}

size_t bbjd_op_bytes_for(const bbjd_pencil_desc *d, int pc, int want_c64)
{
    if (!d || d->n <= 0 || d->bs <= 0) return 0;
    size_t sz[D_END], off[D_END];
    int s = (d->hw > 0 ? d->hw : 1) * d->bs;
    bbjd_int N = (d->n - d->ntau + s - 1) / s;
    return op_layout(d, pc, want_c64, N, s, pc == BBJD_PC_EXACT_LU ? bcr_block_count(N) : 0, sz, off);
}

bbjd_int bbjd_op_size(const bbjd_op *op) { return op ? op->n : 0; }

/* ---- public ---------------------------------------------------------- */
static void op_free_impl(bbjd_op *op)
{
    if (!op) return;
    if (op->dblock) hipFree(op->dblock);
    free(op->even_off); free(op->odd_off);
    free(op);
}

void bbjd_op_free(bbjd_op *op)
{
    if (!op) return;
    const int held = bbjd_devwide_lock();
    op_free_impl(op);
    bbjd_devwide_unlock(held);
}

size_t bbjd_op_bytes(const bbjd_op *op) { return op ? op->dbytes : 0; }

static bbjd_op *op_create_impl(const bbjd_pencil_desc *d, bbjd_zomplex sigma, bbjd_pc_kind pc, int want_c64,
                               void *stream)
{
    if (!d || d->n <= 0 || d->bs <= 0 || d->n % d->bs || d->hw < 0 || d->ntau < 0 || d->ntau > 256 ||
        d->ntau >= d->n)
        return NULL;
    struct bbjd_op *op = calloc(1, sizeof(*op));
    if (!op) return NULL;
    op->n = d->n; op->bs = d->bs; op->hw = d->hw; op->ntau = d->ntau; op->pc = pc; op->want_c64 = want_c64;
    op->nb = d->n / d->bs; op->n1 = d->n - d->ntau; op->nd = 2 * d->hw + 1;
    op->stream = (hipStream_t)stream;
    op->sigma = sigma;
    const bbjd_int n = op->n, n1 = op->n1;
    zcx *bcr_host = NULL;
    size_t bcr_blocks = 0;
    if (pc == BBJD_PC_EXACT_LU) {
        op->s = (d->hw > 0 ? d->hw : 1) * d->bs;
        if (op->s > 64) { op_free_impl(op); return NULL; }
        op->N = (n1 + op->s - 1) / op->s;
        if (bcr_factor(op, d, &bcr_host, &bcr_blocks)) { op_free_impl(op); return NULL; }
    }
    /* device layout */
    size_t off[D_END], sz[D_END];
    const size_t tot = op_layout(d, pc, want_c64, op->N, op->s, bcr_blocks, sz, off);
    const size_t nbandel = (size_t)op->nb * op->nd * op->bs * op->bs;
    const int64_t pe = bbjd_red_part_entries(n, op->ntau > 0 ? op->ntau + 2 : 2);
    if (!HIPOK(hipMalloc(&op->dblock, tot))) { op->dblock = NULL; free(bcr_host); op_free_impl(op); return NULL; }
    op->dbytes = tot;
    char *b = (char *)op->dblock;
#define AT(i) (sz[i] ? (void *)(b + off[i]) : NULL)
    op->band = AT(D_BAND); op->tauc = AT(D_TAUC); op->bdiag = AT(D_BDIAG);
    op->band_c = AT(D_BANDC); op->tauc_c = AT(D_TAUCC); op->bdiag_c = AT(D_BDIAGC);
    op->bj = AT(D_BJ); op->bj_c = AT(D_BJC); op->bcr = AT(D_BCR); op->bvec = AT(D_BV); op->xvec = AT(D_XV);
    op->bvec2 = AT(D_BV2);
    op->Y = AT(D_Y); op->Sinv = AT(D_SINV); op->small = AT(D_SMALL);
    op->rd.part = AT(D_PART); op->rd.pnorm = AT(D_PNORM); op->rd.part_cap = pe;
    {
        char *r = b + off[D_RSC];
        op->rd.scal = (double *)r; op->rd.flag = (int *)(r + 256); op->rd.cnt = (int *)(r + 512);
        op->rd.h = r + 512 + BBJD_CNT_SLOTS * 4; op->rd.h2 = op->rd.h; op->rd.hsum = op->rd.h;
        op->rd.coef = op->rd.h; op->rd.coef2 = op->rd.h;
    }
    op->zt1 = AT(D_ZT1); op->zt2 = AT(D_ZT2); op->doff = AT(D_OFF);
#undef AT
    hipStream_t st = op->stream;
    int rc = 0;
    rc |= !HIPOK(hipMemsetAsync(op->dblock, 0, tot, st));
    rc |= !HIPOK(hipMemcpyAsync(op->band, d->band, nbandel * 16, hipMemcpyHostToDevice, st));
    rc |= !HIPOK(hipMemcpyAsync(op->bdiag, d->bdiag, (size_t)n * 16, hipMemcpyHostToDevice, st));
    rc |= !HIPOK(hipStreamSynchronize(st));   /* later set-up copies use the null stream */
    /* conj(tau), column-major n x ntau (each tau row contiguous) */
    zcx *hb = NULL;
    const size_t hbn = (size_t)n * (op->ntau > 0 ? op->ntau : 1);
    hb = malloc(hbn * sizeof(zcx) > nbandel * sizeof(zcx) ? hbn * sizeof(zcx) : nbandel * sizeof(zcx));
    if (!hb) rc = 1;
    if (!rc && op->ntau > 0) {
        for (size_t i = 0; i < (size_t)n * op->ntau; i++) hb[i] = conj(((const zcx *)d->tau)[i]);
        rc |= !HIPOK(hipMemcpy(op->tauc, hb, (size_t)n * op->ntau * 16, hipMemcpyHostToDevice));
    }
    if (!rc && want_c64) {
        c64 *cb = (c64 *)hb;
        for (size_t i = 0; i < nbandel; i++) {
            cb[i].re = (float)creal(((const zcx *)d->band)[i]);
            cb[i].im = (float)cimag(((const zcx *)d->band)[i]);
        }
        rc |= !HIPOK(hipMemcpy(op->band_c, cb, nbandel * 8, hipMemcpyHostToDevice));
        for (size_t i = 0; i < (size_t)n * op->ntau; i++) {
            const zcx v = conj(((const zcx *)d->tau)[i]);
            cb[i].re = (float)creal(v); cb[i].im = (float)cimag(v);
        }
        if (op->ntau > 0) rc |= !HIPOK(hipMemcpy(op->tauc_c, cb, (size_t)n * op->ntau * 8, hipMemcpyHostToDevice));
        for (bbjd_int i = 0; i < n; i++) {
            cb[i].re = (float)creal(((const zcx *)d->bdiag)[i]);
            cb[i].im = (float)cimag(((const zcx *)d->bdiag)[i]);
        }
        rc |= !HIPOK(hipMemcpy(op->bdiag_c, cb, (size_t)n * 8, hipMemcpyHostToDevice));
    }
    if (!rc && pc == BBJD_PC_BLOCK_JACOBI) {
        const int bs = op->bs;
        const size_t bb = (size_t)bs * bs;
        zcx *inv = malloc((size_t)op->nb * bb * sizeof(zcx));
        if (!inv) rc = 1;
        int bad = 0;
        if (!rc) {
#pragma omp parallel for schedule(static) reduction(| : bad)
            for (bbjd_int bi = 0; bi < op->nb; bi++) {
                zcx *D = inv + (size_t)bi * bb;
                zcx work[64 * 64];
                int piv[64];
                for (int c = 0; c < bs; c++)
                    for (int r = 0; r < bs; r++) D[(size_t)c * bs + r] = m_at(d, sigma, bi * bs + r, bi * bs + c);
                if (bs > 64 || zinv(bs, D, work, piv)) bad |= 1;
            }
            if (bad) rc = 1;
        }
        if (!rc) rc |= !HIPOK(hipMemcpy(op->bj, inv, (size_t)op->nb * bb * 16, hipMemcpyHostToDevice));
        if (!rc && want_c64) {
            c64 *cb = (c64 *)hb;
            c64 *ci = malloc((size_t)op->nb * bb * sizeof(c64));
            if (!ci) rc = 1;
            else {
                for (size_t i = 0; i < (size_t)op->nb * bb; i++) { ci[i].re = (float)creal(inv[i]); ci[i].im = (float)cimag(inv[i]); }
                rc |= !HIPOK(hipMemcpy(op->bj_c, ci, (size_t)op->nb * bb * 8, hipMemcpyHostToDevice));
                free(ci);
            }
            (void)cb;
        }
        free(inv);
    }
    if (!rc && pc == BBJD_PC_EXACT_LU) {
        rc |= !HIPOK(hipMemcpy(op->bcr, bcr_host, bcr_blocks * op->s * op->s * 16, hipMemcpyHostToDevice));
        {
            int64_t offs[2 * 64 + 2];
            memset(offs, 0, sizeof(offs));
            for (int l = 0; l < op->L && l < 64; l++) { offs[2 * l] = op->even_off[l]; offs[2 * l + 1] = op->odd_off[l]; }
            offs[2 * op->L] = op->final_off;
            rc |= !HIPOK(hipMemcpy(op->doff, offs, (size_t)(2 * op->L + 1) * 8, hipMemcpyHostToDevice));
        }
        if (!rc && op->ntau > 0) {
            /* Y = M11^-1 M12 column by column on the device */
            const int nt = op->ntau;
            zcx *col = calloc((size_t)n1, sizeof(zcx));
            zcx *Sh = malloc((size_t)nt * nt * sizeof(zcx));
            zcx *dots = malloc((size_t)nt * sizeof(zcx));
            if (!col || !Sh || !dots) rc = 1;
            for (int a = 0; !rc && a < nt; a++) {
                const bbjd_int lo = n1 - (bbjd_int)(op->hw + 2) * op->bs > 0 ? n1 - (bbjd_int)(op->hw + 2) * op->bs : 0;
                memset(col, 0, (size_t)n1 * sizeof(zcx));
                for (bbjd_int i = lo; i < n1; i++) col[i] = m_at(d, sigma, i, n1 + a);
                rc |= !HIPOK(hipMemcpy(op->bvec, col, (size_t)n1 * 16, hipMemcpyHostToDevice));
                rc |= bcr_solve(op, st, 1, NULL, NULL, 0, NULL, 0);
                rc |= !HIPOK(hipMemcpyAsync((zcx *)op->Y + (size_t)a * n1, op->xvec, (size_t)n1 * 16,
                                            hipMemcpyDeviceToDevice, st));
                /* M21 Y[:, a] */
                rc |= bbjd_k_gemv_c(st, n1, nt, op->tauc, n, (zcx *)op->Y + (size_t)a * n1, NULL, op->small, NULL,
                                    &op->rd);
                rc |= !HIPOK(hipMemcpyAsync(dots, op->small, (size_t)nt * 16, hipMemcpyDeviceToHost, st));
                rc |= !HIPOK(hipStreamSynchronize(st));
                for (int t = 0; t < nt; t++) Sh[(size_t)a * nt + t] = m_at(d, sigma, n1 + t, n1 + a) - dots[t];
            }
            if (!rc) {
                zcx *work = malloc((size_t)nt * nt * sizeof(zcx));
                int *piv = malloc((size_t)nt * sizeof(int));
                if (!work || !piv || zinv(nt, Sh, work, piv)) rc = 1;
                else rc |= !HIPOK(hipMemcpy(op->Sinv, Sh, (size_t)nt * nt * 16, hipMemcpyHostToDevice));
                free(work); free(piv);
            }
            /* the padded RHS entries of bvec must be zero for every apply */
            rc |= !HIPOK(hipMemsetAsync(op->bvec, 0, (size_t)2 * op->N * op->s * 16, st));
            free(col); free(Sh); free(dots);
        }
    }
    free(hb);
    free(bcr_host);
    rc |= !HIPOK(hipStreamSynchronize(st));
    if (rc) { op_free_impl(op); return NULL; }
    return op;
}

bbjd_op *bbjd_op_create(const bbjd_pencil_desc *d, bbjd_zomplex sigma, bbjd_pc_kind pc, int want_c64, void *stream)
{
    const int held = bbjd_devwide_lock();
    bbjd_op *op = op_create_impl(d, sigma, pc, want_c64, stream);
    bbjd_devwide_unlock(held);
    return op;
}

int bbjd_op_matvec_A(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *data)
{
    struct bbjd_op *op = (struct bbjd_op *)data;
    if (!op || n != op->n) return -1;
    hipStream_t st = op_stream(op);
    int rc = bbjd_o_band_mv(st, op->n1, op->nb, op->bs, op->hw, op->band, x, y, 0);
    if (op->ntau > 0)
        rc |= bbjd_k_gemv_c(st, n, op->ntau, op->tauc, n, x, NULL, (zcx *)y + op->n1, NULL, &op->rd);
    return rc ? -1 : 0;
}

int bbjd_op_matvec_AB(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *ya, bbjd_zomplex *yb, void *data)
{
    struct bbjd_op *op = (struct bbjd_op *)data;
    if (!op || n != op->n) return -1;
    hipStream_t st = op_stream(op);
    const int e = bbjd_o_band_ab(st, n, op->n1, op->nb, op->bs, op->hw, op->band, op->bdiag, op->ntau, op->tauc, x,
                                 ya, yb, op->rd.part, op->rd.part_cap, 0);
    if (e != -3) return e ? -1 : 0;
    if (bbjd_op_matvec_A(n, x, ya, data)) return -1;
    return bbjd_op_matvec_B(n, x, yb, data);
}

int bbjd_op_matvec_ABr(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *ya, bbjd_zomplex *yb, bbjd_zomplex alpha,
                       bbjd_zomplex beta, bbjd_zomplex *r, double *rnorm, void *data)
{
    struct bbjd_op *op = (struct bbjd_op *)data;
    if (!op || n != op->n || !r || !rnorm) return -1;
    hipStream_t st = op_stream(op);
    const int64_t pn = (n + BBJD_TILE_MIN_ROWS - 1) / BBJD_TILE_MIN_ROWS;
    const int e = bbjd_o_band_abr(st, n, op->n1, op->nb, op->bs, op->hw, op->band, op->bdiag, op->ntau, op->tauc, x,
                                  ya, yb, op->rd.part, op->rd.part_cap, 0, creal(alpha), cimag(alpha), creal(beta),
                                  cimag(beta), r, op->rd.pnorm, pn, rnorm);
    if (e != -3) return e ? -1 : 0;
    /* shapes the fused pass does not take: the two matvecs, then r and its
     * norm through the operator's own scalar slot */
    if (bbjd_op_matvec_AB(n, x, ya, yb, data)) return -1;
    int rc = bbjd_k_lincomb_norm(st, n, creal(alpha), cimag(alpha), ya, creal(beta), cimag(beta), yb, r, &op->rd,
                                 SL_T0);
    rc |= !HIPOK(hipMemcpyAsync(rnorm, op->rd.scal + SL_T0, sizeof(double), hipMemcpyDeviceToDevice, st));
    return rc ? -1 : 0;
}

int bbjd_op_matvec_B(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *data)
{
    struct bbjd_op *op = (struct bbjd_op *)data;
    if (!op || n != op->n) return -1;
    return bbjd_o_diag_mv(op_stream(op), n, op->bdiag, x, y, 0) ? -1 : 0;
}

int bbjd_op_precond(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *data)
{
    struct bbjd_op *op = (struct bbjd_op *)data;
    if (!op || n != op->n) return -1;
    hipStream_t st = op_stream(op);
    switch (op->pc) {
    case BBJD_PC_EXACT_LU: return exact_apply(op, st, 1, x, y, NULL, NULL, 0) ? -1 : 0;
    case BBJD_PC_BLOCK_JACOBI: return bbjd_o_bj_apply(st, n, op->bs, op->bj, x, y, 0) ? -1 : 0;
    default: return bbjd_copy_async(y, x, (size_t)n * 16, hipMemcpyDeviceToDevice, st) ? -1 : 0;
    }
}

int bbjd_op_precond2(bbjd_int n, const bbjd_zomplex *x0, bbjd_zomplex *y0, const bbjd_zomplex *x1,
                     bbjd_zomplex *y1, void *data)
{
    struct bbjd_op *op = (struct bbjd_op *)data;
    if (!op || n != op->n) return -1;
    if (op->pc == BBJD_PC_EXACT_LU) return exact_apply(op, op_stream(op), 2, x0, y0, x1, y1, 0) ? -1 : 0;
    if (bbjd_op_precond(n, x0, y0, data)) return -1;
    return bbjd_op_precond(n, x1, y1, data);
}

int bbjd_op_matvec_A_c64(bbjd_int n, const bbjd_cfloat *x, bbjd_cfloat *y, void *data)
{
    struct bbjd_op *op = (struct bbjd_op *)data;
    if (!op || n != op->n || !op->want_c64) return -1;
    hipStream_t st = op_stream(op);
    int rc = bbjd_o_band_mv(st, op->n1, op->nb, op->bs, op->hw, op->band_c, x, y, 1);
    if (op->ntau > 0)
        rc |= bbjd_k_gemv_c_c(st, n, op->ntau, op->tauc_c, n, x, NULL, (c64 *)y + op->n1, NULL, &op->rd);
    return rc ? -1 : 0;
}

int bbjd_op_matvec_AB_c64(bbjd_int n, const bbjd_cfloat *x, bbjd_cfloat *ya, bbjd_cfloat *yb, void *data)
{
    struct bbjd_op *op = (struct bbjd_op *)data;
    if (!op || n != op->n || !op->want_c64) return -1;
    hipStream_t st = op_stream(op);
    const int e = bbjd_o_band_ab(st, n, op->n1, op->nb, op->bs, op->hw, op->band_c, op->bdiag_c, op->ntau,
                                 op->tauc_c, x, ya, yb, op->rd.part, op->rd.part_cap, 1);
    if (e != -3) return e ? -1 : 0;
    if (bbjd_op_matvec_A_c64(n, x, ya, data)) return -1;
    return bbjd_op_matvec_B_c64(n, x, yb, data);
}

int bbjd_op_matvec_B_c64(bbjd_int n, const bbjd_cfloat *x, bbjd_cfloat *y, void *data)
{
    struct bbjd_op *op = (struct bbjd_op *)data;
    if (!op || n != op->n || !op->want_c64) return -1;
    return bbjd_o_diag_mv(op_stream(op), n, op->bdiag_c, x, y, 1) ? -1 : 0;
}

int bbjd_op_precond_c64(bbjd_int n, const bbjd_cfloat *x, bbjd_cfloat *y, void *data)
{
    struct bbjd_op *op = (struct bbjd_op *)data;
    if (!op || n != op->n || !op->want_c64) return -1;
    hipStream_t st = op_stream(op);
    switch (op->pc) {
    case BBJD_PC_BLOCK_JACOBI: return bbjd_o_bj_apply(st, n, op->bs, op->bj_c, x, y, 1) ? -1 : 0;
    case BBJD_PC_EXACT_LU: {
        /* the FP64 solve reading and writing complex64 directly: the same
         * values as widening, solving and rounding in three passes */
        if (g_c64_io) {
            const int e = exact_apply(op, st, 1, x, y, NULL, NULL, 1);
            if (e != -3) return e ? -1 : 0;
        }
        int rc = bbjd_k_c2z(st, n, x, op->zt1);
        rc |= exact_apply(op, st, 1, op->zt1, op->zt2, NULL, NULL, 0);
        rc |= bbjd_k_z2c(st, n, op->zt2, y);
        return rc ? -1 : 0;
    }
    default: return bbjd_copy_async(y, x, (size_t)n * 8, hipMemcpyDeviceToDevice, st) ? -1 : 0;
    }
}
