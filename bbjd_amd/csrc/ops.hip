// ops.hip -- device kernels of the block-banded + tau-row operators.
//
//  band matvec : one thread per row; the row's 2hw+1 blocks are read
//                column-major so lanes of a block row read consecutive
//                addresses; x is read through L1/L2 by all bs rows.
//  diag / block-Jacobi applies : elementwise / bs x bs per block row.
//  block cyclic reduction (BCR) solve of the block-tridiagonal leading part
//                of A - sigma B: log2(N) forward levels b_e -= alpha b_{e-h}
//                + gamma b_{e+h} over the surviving rows, one final s x s
//                solve, log2(N) backward levels x_o = Dinv b_o - P x_{o-h}
//                - Q x_{o+h}; every level is one launch with one thread per
//                (block row, row) and no intra-level dependencies.
//  border      : the tau rows through the Schur complement.
// Templated on the real type for the complex64 copies of the mixed path.
#include <hip/hip_runtime.h>
#include "ops_dev.h"
#include "launch.h"

typedef long long i64;
template <typename R> struct alignas(2 * sizeof(R)) Cx { R re, im; };
// a += q h and a -= q h as four explicit fused multiply-adds into a: the
// compiler has no choice of which products to fuse, so kernels that apply
// them in the same order round identically whatever the surrounding code
__device__ __forceinline__ double ffma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float ffma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <typename R> __device__ __forceinline__ void cfma(Cx<R> &a, Cx<R> q, Cx<R> h)
{
    a.re = ffma(q.re, h.re, a.re);
    a.re = ffma(-q.im, h.im, a.re);
    a.im = ffma(q.re, h.im, a.im);
    a.im = ffma(q.im, h.re, a.im);
}
__device__ __forceinline__ Cx<double> cx(double a, double b) { Cx<double> c; c.re = a; c.im = b; return c; }
__device__ __forceinline__ Cx<double> cx0() { return cx(0.0, 0.0); }
template <typename R> __device__ __forceinline__ void cfms(Cx<R> &a, Cx<R> q, Cx<R> h)
{
    a.re = ffma(-q.re, h.re, a.re);
    a.re = ffma(q.im, h.im, a.re);
    a.im = ffma(-q.re, h.im, a.im);
    a.im = ffma(-q.im, h.re, a.im);
}

static constexpr int TPB = 256;

template <typename R>
__device__ __forceinline__ void k_band_mv_b(const VB vb, i64 n1, i64 nb, int bs, int hw, const Cx<R> *__restrict__ band,
                                                 const Cx<R> *__restrict__ x, Cx<R> *__restrict__ y)
{
    const i64 i = (i64)vb.bx * TPB + threadIdx.x;
    if (i >= n1) return;
    const i64 bi = i / bs;
    const int r = (int)(i % bs), nd = 2 * hw + 1;
    Cx<R> acc = {0, 0};
    for (int d = 0; d < nd; d++) {
        const i64 bj = bi + d - hw;
        if (bj < 0 || bj >= nb) continue;
        const Cx<R> *blk = band + (bi * nd + d) * bs * bs + r;
        const Cx<R> *xs = x + bj * bs;
        for (int c = 0; c < bs; c++) cfma(acc, blk[(i64)c * bs], xs[c]);
    }
    y[i] = acc;
}
template <typename R>
__global__ __launch_bounds__(TPB) void k_band_mv(i64 n1, i64 nb, int bs, int hw, const Cx<R> *__restrict__ band,
                                                 const Cx<R> *__restrict__ x, Cx<R> *__restrict__ y)
{
    k_band_mv_b<R>(VB_HW, n1, nb, bs, hw, band, x, y);
}

template <typename R>
__device__ __forceinline__ void k_diag_mv_b(const VB vb, i64 n, const Cx<R> *__restrict__ dg,
                                            const Cx<R> *__restrict__ x,
                                                 Cx<R> *__restrict__ y)
{
    for (i64 i = (i64)vb.bx * TPB + threadIdx.x; i < n; i += (i64)vb.gx * TPB) {
        Cx<R> a = {0, 0};
        cfma(a, dg[i], x[i]);
        y[i] = a;
    }
}
template <typename R>
__global__ __launch_bounds__(TPB) void k_diag_mv(i64 n, const Cx<R> *__restrict__ dg, const Cx<R> *__restrict__ x,
                                                 Cx<R> *__restrict__ y)
{
    k_diag_mv_b<R>(VB_HW, n, dg, x, y);
}

// y_A = A x and y_B = B x in one pass over x (the block-banded rows i < n1,
// the diagonal of B on every row, and per-block partial dots of the ntau
// dense rows, conj(tauc[:, c]) . x, finished by k_tau_finish).  The band
// rows and the diagonal are k_band_mv's and k_diag_mv's arithmetic.  With r
// non-NULL also r = a y_A + b y_B on the band rows (the engine's k_lincomb
// arithmetic) and per-block partial sums of |r|^2 in rpart; the dense rows'
// r and the norm come in k_tau_finish.
template <typename R>
__device__ __forceinline__ Cx<R> wave_sum(Cx<R> v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        v.re += __shfl_xor(v.re, o, 64);
        v.im += __shfl_xor(v.im, o, 64);
    }
    return v;
}

template <typename R>
__device__ __forceinline__ R wave_sum_r(R v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// a y_A + b y_B as the engine's k_lincomb forms it: (a * ya) + (b * yb)
template <typename R>
__device__ __forceinline__ Cx<R> lin2(Cx<R> a, Cx<R> x, Cx<R> b, Cx<R> y)
{
    const Cx<R> p = {a.re * x.re - a.im * x.im, a.re * x.im + a.im * x.re};
    const Cx<R> q = {b.re * y.re - b.im * y.im, b.re * y.im + b.im * y.re};
    return Cx<R>{p.re + q.re, p.im + q.im};
}

static constexpr int AB_MAX_TAU = 16;

// BS, HW > 0: the block size and half-width at compile time, so a row's
// (2 HW + 1) BS band loads are all in flight at once; 0: runtime bs, hw.
// The dense rows' values load four at a time.
template <typename R, int BS, int HW>
__device__ __forceinline__ void k_band_ab_b(const VB vb, i64 n, i64 n1, i64 nb, int bs_rt, int hw_rt,
                                                 const Cx<R> *__restrict__ band, const Cx<R> *__restrict__ dg,
                                                 int ntau, const Cx<R> *__restrict__ tauc,
                                                 const Cx<R> *__restrict__ x, Cx<R> *__restrict__ ya,
                                                 Cx<R> *__restrict__ yb, Cx<R> *__restrict__ part, Cx<R> ca,
                                                 Cx<R> cb, Cx<R> *__restrict__ r, R *__restrict__ rpart)
{
    __shared__ Cx<R> red[AB_MAX_TAU][TPB / 64];
    __shared__ R rred[TPB / 64];
    const i64 i = (i64)vb.bx * TPB + threadIdx.x;
    Cx<R> xi = {0, 0};
    R rs = 0;
    if (i < n) {
        xi = x[i];
        Cx<R> b = {0, 0};
        cfma(b, dg[i], xi);
        yb[i] = b;
        if (i < n1) {
            const int bs = BS > 0 ? BS : bs_rt, hw = BS > 0 ? HW : hw_rt;
            const i64 bi = i / bs;
            const int rr = (int)(i % bs), nd = 2 * hw + 1;
            Cx<R> acc = {0, 0};
            if constexpr (BS > 0) {
#pragma unroll
                for (int d = 0; d < 2 * HW + 1; d++) {
                    const i64 bj = bi + d - HW;
                    if (bj < 0 || bj >= nb) continue;
                    const Cx<R> *blk = band + (bi * nd + d) * BS * BS + rr;
                    const Cx<R> *xs = x + bj * BS;
#pragma unroll
                    for (int c = 0; c < BS; c++) cfma(acc, blk[(i64)c * BS], xs[c]);
                }
            } else {
                for (int d = 0; d < nd; d++) {
                    const i64 bj = bi + d - hw;
                    if (bj < 0 || bj >= nb) continue;
                    const Cx<R> *blk = band + (bi * nd + d) * bs * bs + rr;
                    const Cx<R> *xs = x + bj * bs;
                    for (int c = 0; c < bs; c++) cfma(acc, blk[(i64)c * bs], xs[c]);
                }
            }
            ya[i] = acc;
            if (r) {
                const Cx<R> v = lin2(ca, acc, cb, b);
                r[i] = v;
                rs = v.re * v.re + v.im * v.im;
            }
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c0 = 0; c0 < ntau; c0 += 4) {
        Cx<R> q[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            q[k] = Cx<R>{0, 0};
            if (i < n && c0 + k < ntau) q[k] = tauc[i + (i64)(c0 + k) * n];
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (c0 + k >= ntau) break;
            Cx<R> p = {0, 0};
            if (i < n) {
                // conj(q) x
                p.re = q[k].re * xi.re + q[k].im * xi.im;
                p.im = q[k].re * xi.im - q[k].im * xi.re;
            }
            p = wave_sum(p);
            if (lane == 0) red[c0 + k][w] = p;
        }
    }
    if (r) {
        rs = wave_sum_r(rs);
        if (lane == 0) rred[w] = rs;
    }
    __syncthreads();
    if (threadIdx.x < ntau) {
        Cx<R> s = red[threadIdx.x][0];
#pragma unroll
        for (int v = 1; v < TPB / 64; v++) { s.re += red[threadIdx.x][v].re; s.im += red[threadIdx.x][v].im; }
        part[(i64)vb.bx * ntau + threadIdx.x] = s;
    }
    if (r && threadIdx.x == 0) {
        R s = rred[0];
#pragma unroll
        for (int v = 1; v < TPB / 64; v++) s += rred[v];
        rpart[vb.bx] = s;
    }
}
template <typename R, int BS, int HW>
__global__ __launch_bounds__(TPB) void k_band_ab(i64 n, i64 n1, i64 nb, int bs_rt, int hw_rt,
                                                 const Cx<R> *__restrict__ band, const Cx<R> *__restrict__ dg,
                                                 int ntau, const Cx<R> *__restrict__ tauc,
                                                 const Cx<R> *__restrict__ x, Cx<R> *__restrict__ ya,
                                                 Cx<R> *__restrict__ yb, Cx<R> *__restrict__ part, Cx<R> ca,
                                                 Cx<R> cb, Cx<R> *__restrict__ r, R *__restrict__ rpart)
{
    k_band_ab_b<R, BS, HW>(VB_HW, n, n1, nb, bs_rt, hw_rt, band, dg, ntau, tauc, x, ya, yb, part, ca, cb, r, rpart);
}

// y_A's dense rows: wave c sums column c of the nblk block partials in a
// fixed order.  With r: wave ntau sums the band rows' |r|^2 partials, then
// forms the dense rows' r = a y_A + b y_B and writes ||r|| to *rnorm.
template <typename R>
__device__ __forceinline__ void k_tau_finish_b(const VB vb, i64 nblk, int ntau, i64 n1,
                                                                      const Cx<R> *__restrict__ part,
                                                                      Cx<R> *__restrict__ ya,
                                                                      const Cx<R> *__restrict__ yb, Cx<R> ca,
                                                                      Cx<R> cb, Cx<R> *__restrict__ r,
                                                                      const R *__restrict__ rpart,
                                                                      double *__restrict__ rnorm)
{
    __shared__ Cx<R> tv[AB_MAX_TAU];
    const int lane = threadIdx.x & 63, c = threadIdx.x >> 6;
    R rs = 0;
    if (c < ntau) {
        Cx<R> s = {0, 0};
        for (i64 b = lane; b < nblk; b += 64) {
            const Cx<R> v = part[b * ntau + c];
            s.re += v.re; s.im += v.im;
        }
        s = wave_sum(s);
        if (lane == 0) { ya[n1 + c] = s; tv[c] = s; }
    } else if (c == ntau && r) {
        for (i64 b = lane; b < nblk; b += 64) rs += rpart[b];
        rs = wave_sum_r(rs);
    }
    __syncthreads();
    if (r && c == ntau && lane == 0) {
        for (int t = 0; t < ntau; t++) {
            const Cx<R> v = lin2(ca, tv[t], cb, yb[n1 + t]);
            r[n1 + t] = v;
            rs += v.re * v.re + v.im * v.im;
        }
        *rnorm = sqrt((double)rs);
    }
}
template <typename R>
__global__ __launch_bounds__(64 * AB_MAX_TAU) void k_tau_finish(i64 nblk, int ntau, i64 n1,
                                                                      const Cx<R> *__restrict__ part,
                                                                      Cx<R> *__restrict__ ya,
                                                                      const Cx<R> *__restrict__ yb, Cx<R> ca,
                                                                      Cx<R> cb, Cx<R> *__restrict__ r,
                                                                      const R *__restrict__ rpart,
                                                                      double *__restrict__ rnorm)
{
    k_tau_finish_b<R>(VB_HW, nblk, ntau, n1, part, ya, yb, ca, cb, r, rpart, rnorm);
}

template <typename R>
__device__ __forceinline__ void k_bj_apply_b(const VB vb, i64 n, int bs, const Cx<R> *__restrict__ inv,
                                                  const Cx<R> *__restrict__ x, Cx<R> *__restrict__ y)
{
    const i64 i = (i64)vb.bx * TPB + threadIdx.x;
    if (i >= n) return;
    const i64 bi = i / bs;
    const int r = (int)(i % bs);
    const Cx<R> *D = inv + bi * bs * bs + r;
    const Cx<R> *xs = x + bi * bs;
    Cx<R> a = {0, 0};
    for (int c = 0; c < bs; c++) cfma(a, D[(i64)c * bs], xs[c]);
    y[i] = a;
}
template <typename R>
__global__ __launch_bounds__(TPB) void k_bj_apply(i64 n, int bs, const Cx<R> *__restrict__ inv,
                                                  const Cx<R> *__restrict__ x, Cx<R> *__restrict__ y)
{
    k_bj_apply_b<R>(VB_HW, n, bs, inv, x, y);
}

// BCR on NR right-hand sides at once (vector q at b + q vs): the factor
// blocks are read once for all of them.
//
// forward level: evens m = 2j at stride h, j < ne:
//   b_i -= alpha_j b_{i-h} + gamma_j b_{i+h}
template <int NR, int S>
__device__ __forceinline__ void bcr_fwd_row(int s_rt, i64 M, i64 h, i64 j, int r, const Cx<double> *__restrict__ AG,
                                            Cx<double> *b, i64 vs)
{
    const int s = S > 0 ? S : s_rt;
    const i64 m = 2 * j, i = m * h, ss = (i64)s * s;
    Cx<double> acc[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) acc[q] = b[q * vs + i * s + r];
    if (m >= 1) {
        const Cx<double> *al = AG + j * 2 * ss + r;
        const i64 o = (i - h) * s;
#pragma unroll
        for (int c = 0; c < s; c++) {
            const Cx<double> a = al[(i64)c * s];
#pragma unroll
            for (int q = 0; q < NR; q++) cfms(acc[q], a, b[q * vs + o + c]);
        }
    }
    if (m + 1 < M) {
        const Cx<double> *ga = AG + j * 2 * ss + ss + r;
        const i64 o = (i + h) * s;
#pragma unroll
        for (int c = 0; c < s; c++) {
            const Cx<double> g = ga[(i64)c * s];
#pragma unroll
            for (int q = 0; q < NR; q++) cfms(acc[q], g, b[q * vs + o + c]);
        }
    }
#pragma unroll
    for (int q = 0; q < NR; q++) b[q * vs + i * s + r] = acc[q];
}

// backward level: odds m = 2j + 1, DPQ = (Dinv, Dinv L, Dinv U):
//   x_i = Dinv b_i - P x_{i-h} - Q x_{i+h}
template <int NR, int S>
__device__ __forceinline__ void bcr_bwd_row(int s_rt, i64 M, i64 h, i64 j, int r, const Cx<double> *__restrict__ DPQ,
                                            const Cx<double> *b, Cx<double> *x, i64 vs)
{
    const int s = S > 0 ? S : s_rt;
    const i64 m = 2 * j + 1, i = m * h, ss = (i64)s * s;
    const Cx<double> *D = DPQ + j * 3 * ss + r, *P = D + ss, *Qm = P + ss;
    Cx<double> acc[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) acc[q] = cx0();
    const i64 ob = i * s, ol = (i - h) * s;
#pragma unroll
    for (int c = 0; c < s; c++) {
        const Cx<double> dd = D[(i64)c * s], pp = P[(i64)c * s];
#pragma unroll
        for (int q = 0; q < NR; q++) { cfma(acc[q], dd, b[q * vs + ob + c]); cfms(acc[q], pp, x[q * vs + ol + c]); }
    }
    if (m + 1 < M) {
        const i64 orr = (i + h) * s;
#pragma unroll
        for (int c = 0; c < s; c++) {
            const Cx<double> qq = Qm[(i64)c * s];
#pragma unroll
            for (int q = 0; q < NR; q++) cfms(acc[q], qq, x[q * vs + orr + c]);
        }
    }
#pragma unroll
    for (int q = 0; q < NR; q++) x[q * vs + i * s + r] = acc[q];
}

// Level 0 of the forward reduction reading the right-hand sides from the
// caller's vectors (n1 rows; the padding rows of the last super-block read
// as zero): every row of b is written, the even super-rows reduced
// (b_i = x_i - alpha_j x_{i-1} - gamma_j x_{i+1}), the odd ones copied --
// the copy of the input into b and the first level in one pass.
template <int NR, int S>
__global__ __launch_bounds__(TPB) void k_bcr_fwd0(int s_rt, i64 N, i64 n1, const Cx<double> *__restrict__ AG,
                                                  const Cx<double> *__restrict__ x0,
                                                  const Cx<double> *__restrict__ x1, Cx<double> *b, i64 vs)
{
    const int s = S > 0 ? S : s_rt;
    const i64 gid = (i64)blockIdx.x * TPB + threadIdx.x;
    const i64 i = gid / s, ss = (i64)s * s;
    const int r = (int)(gid % s);
    if (i >= N) return;
    const Cx<double> *xs[2] = {x0, x1};
    Cx<double> acc[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) {
        const i64 e = i * s + r;
        acc[q] = e < n1 ? xs[q][e] : cx0();
    }
    if ((i & 1) == 0) {
        const i64 j = i / 2;
        // neighbour super-row (i -/+ 1) times its coupling block; the left
        // neighbour is always whole, the right one may hold padding rows
        auto sub = [&](const Cx<double> *blk, i64 o) {
            if (o + s <= n1) {
#pragma unroll
                for (int c = 0; c < s; c++) {
                    const Cx<double> a = blk[(i64)c * s];
#pragma unroll
                    for (int q = 0; q < NR; q++) cfms(acc[q], a, xs[q][o + c]);
                }
            } else {
                for (int c = 0; c < s && o + c < n1; c++) {
                    const Cx<double> a = blk[(i64)c * s];
#pragma unroll
                    for (int q = 0; q < NR; q++) cfms(acc[q], a, xs[q][o + c]);
                }
            }
        };
        if (i >= 1) sub(AG + j * 2 * ss + r, (i - 1) * s);
        if (i + 1 < N) sub(AG + j * 2 * ss + ss + r, (i + 1) * s);
    }
#pragma unroll
    for (int q = 0; q < NR; q++) b[q * vs + i * s + r] = acc[q];
}

// S: the super-block size when known at compile time (the row loops unroll
// and every factor load of a row is in flight at once), 0 for any s.
template <int NR, int S>
__global__ __launch_bounds__(TPB) void k_bcr_fwd(int s, i64 M, i64 h, i64 ne, const Cx<double> *__restrict__ AG,
                                                 Cx<double> *b, i64 vs)
{
    const int sv = S > 0 ? S : s;
    const i64 gid = (i64)blockIdx.x * TPB + threadIdx.x;
    const i64 j = gid / sv;
    if (j >= ne) return;
    bcr_fwd_row<NR, S>(s, M, h, j, (int)(gid % sv), AG, b, vs);
}

template <int NR, int S>
__global__ __launch_bounds__(TPB) void k_bcr_bwd(int s, i64 M, i64 h, i64 no, const Cx<double> *__restrict__ DPQ,
                                                 const Cx<double> *b, Cx<double> *x, i64 vs)
{
    const int sv = S > 0 ? S : s;
    const i64 gid = (i64)blockIdx.x * TPB + threadIdx.x;
    const i64 j = gid / sv;
    if (j >= no) return;
    bcr_bwd_row<NR, S>(s, M, h, j, (int)(gid % sv), DPQ, b, x, vs);
}

// G backward levels l + G - 1 .. l in one launch.  Block rows at multiples
// of H = 2^(l+G) come from deeper levels; each block owns U units of H
// block rows, holds its units' rows in LDS (slot = row offset / 2^l from
// the first unit's start, slot U 2^G the next block's first boundary) and
// solves them one level per phase, top level first: phase g solves the odd
// rows of level l + g, whose neighbours at +- 2^(l+g) are boundaries or rows
// of an earlier phase.  Each row's arithmetic is bcr_bwd_row's, in the same
// order.  U fills the last phase's TPB threads (2^(G-1) rows per unit).
// bl.f[g]: the buffer holding level l + g's right-hand sides.
struct BcrLevels { const Cx<double> *f[4]; };

template <int NR, int S, int G>
__global__ __launch_bounds__(TPB) void k_bcr_bwd_multi(i64 N, int l, BcrLevels dpq, BcrLevels bl, Cx<double> *x,
                                                       i64 vs, const Cx<double> *__restrict__ tauc, int ntau, i64 n1,
                                                       i64 ldt, Cx<double> *__restrict__ dpart)
{
    constexpr int U = TPB / ((1 << (G - 1)) * S);
    constexpr int SLOTS = U * (1 << G) + 1;
    constexpr i64 ss = (i64)S * S;
    __shared__ Cx<double> lx[NR][SLOTS][S];
    const int t = threadIdx.x;
    const i64 u0 = (i64)blockIdx.x * U;
    for (int e = t; e < (U + 1) * S * NR; e += TPB) {
        const int q = e / ((U + 1) * S), rem = e % ((U + 1) * S), u = rem / S, r = rem % S;
        const i64 i = ((u0 + u) << G) << l;
        lx[q][u << G][r] = i < N ? x[q * vs + i * S + r] : cx0();
    }
    __syncthreads();
#pragma unroll
    for (int g = G - 1; g >= 0; g--) {
        const int R = 1 << (G - 1 - g);
        const int lv = l + g;
        const i64 hg = (i64)1 << lv, M = (N + hg - 1) >> lv;
        if (t < U * R * S) {
            const int rr = t / S, r = t % S, u = rr / R, k = rr % R;
            const int slot = (u << G) + ((2 * k + 1) << g);
            const i64 m = ((u0 + u) << (G - g)) + 2 * k + 1;
            if (m < M) {
                const i64 i = m * hg, j = (m - 1) / 2;
                const Cx<double> *D = dpq.f[g] + j * 3 * ss + r, *P = D + ss, *Qm = P + ss;
                const Cx<double> *b = bl.f[g];
                Cx<double> acc[NR];
#pragma unroll
                for (int q = 0; q < NR; q++) acc[q] = cx0();
                const i64 ob = i * S;
#pragma unroll
                for (int c = 0; c < S; c++) {
                    const Cx<double> dd = D[(i64)c * S], pp = P[(i64)c * S];
#pragma unroll
                    for (int q = 0; q < NR; q++) {
                        cfma(acc[q], dd, b[q * vs + ob + c]);
                        cfms(acc[q], pp, lx[q][slot - (1 << g)][c]);
                    }
                }
                if (m + 1 < M) {
#pragma unroll
                    for (int c = 0; c < S; c++) {
                        const Cx<double> qq = Qm[(i64)c * S];
#pragma unroll
                        for (int q = 0; q < NR; q++) cfms(acc[q], qq, lx[q][slot + (1 << g)][c]);
                    }
                }
#pragma unroll
                for (int q = 0; q < NR; q++) {
                    x[q * vs + ob + r] = acc[q];
                    lx[q][slot][r] = acc[q];
                }
            }
        }
        __syncthreads();
    }
    // with tauc (the group that ends at level 0, where the block's LDS holds
    // every row of its units): partial dots conj(tauc[:, t]) . x over the
    // block's rows below n1, per (t, q), for the dense rows' border
    if (tauc) {
        __shared__ Cx<double> dred[16][NR][TPB / 64];
        const int lane = t & 63, w = t >> 6;
        for (int tt = 0; tt < ntau; tt++) {
            Cx<double> p[NR];
#pragma unroll
            for (int q = 0; q < NR; q++) p[q] = cx0();
            for (int e = t; e < U * (1 << G) * S; e += TPB) {
                const int slot = e / S, r = e % S;
                const i64 el = ((u0 << G) + slot) * S + r;
                if (el < n1 && ((u0 << G) + slot) < N) {
                    const Cx<double> tq = tauc[el + (i64)tt * ldt];
#pragma unroll
                    for (int q = 0; q < NR; q++) {
                        const Cx<double> v = lx[q][slot][r];
                        p[q].re = ffma(tq.re, v.re, p[q].re);
                        p[q].re = ffma(tq.im, v.im, p[q].re);
                        p[q].im = ffma(tq.re, v.im, p[q].im);
                        p[q].im = ffma(-tq.im, v.re, p[q].im);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NR; q++) {
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) {
                    p[q].re += __shfl_xor(p[q].re, o, 64);
                    p[q].im += __shfl_xor(p[q].im, o, 64);
                }
                if (lane == 0) dred[tt][q][w] = p[q];
            }
        }
        __syncthreads();
        if (t < ntau * NR) {
            const int tt = t / NR, q = t % NR;
            Cx<double> s = dred[tt][q][0];
#pragma unroll
            for (int v = 1; v < TPB / 64; v++) { s.re += dred[tt][q][v].re; s.im += dred[tt][q][v].im; }
            dpart[(i64)blockIdx.x * ntau * NR + t] = s;
        }
    }
}

// Forward levels l .. l + G - 1 in one launch.  Rows are those at
// multiples of 2^l ("level-l rows", index i, block row i 2^l), grouped in
// units of 2^G; each block reduces its U units plus the unit to their left
// (rows 1 .. 2^G - 1 of it reach its first unit within G levels) in LDS, in
// place, level by level, and writes its own units' level-l rows to bout --
// every one of them final for the levels below l + G (odd rows of level
// l + g keep their value after level l + g - 1; the unit boundaries carry
// level l + G - 1's update).  The input is never written by the launch, so
// the recomputed halo is exact: with x0 (x1), l = 0 and the caller's n1-row
// right-hand sides (padding rows read as zero); else bin, another buffer
// than bout.  Each row's arithmetic is k_bcr_fwd0's and bcr_fwd_row's.  U + 1
// units make 2 TPB / S rows, so level l's TPB / S - 1 even rows fill the block.
template <int NR, int S, int G>
__global__ __launch_bounds__(TPB) void k_bcr_fwd_multi(i64 N, int l, i64 n1, BcrLevels ag,
                                                       const Cx<double> *__restrict__ x0,
                                                       const Cx<double> *__restrict__ x1,
                                                       const Cx<double> *__restrict__ bin, Cx<double> *bout, i64 vs)
{
    constexpr int U = TPB / ((1 << (G - 1)) * S) - 1;
    constexpr int ROWS = (U + 1) << G;          // halo unit + own units: 2 TPB / S
    constexpr i64 ss = (i64)S * S;
    __shared__ Cx<double> lb[NR][ROWS][S];
    const int t = threadIdx.x;
    const i64 p0 = ((i64)blockIdx.x * U - 1) << G;   // level-l row of slot 0 (the halo unit)
    const Cx<double> *xs[2] = {x0, x1};
    for (int e = t; e < ROWS * S * NR; e += TPB) {
        const int q = e / (ROWS * S), rem = e % (ROWS * S), sl = rem / S, r = rem % S;
        const i64 il = p0 + sl, bi = il << l;
        Cx<double> v = cx0();
        if (il >= 0 && bi < N) {
            if (x0) {
                const i64 el = bi * S + r;
                if (el < n1) v = xs[q][el];
            } else
                v = bin[q * vs + bi * S + r];
        }
        lb[q][sl][r] = v;
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int lv = l + g;
        const i64 h = (i64)1 << g, M = (N + ((i64)1 << lv) - 1) >> lv;
        // even rows of level l + g: slots that are multiples of 2^(g+1),
        // except slot 0 (no row of the block needs it); within a level, rows
        // read only their own value and odd neighbours, which the level leaves
        const int nrow = (ROWS >> (g + 1)) - 1;
        for (int e = t; e < nrow * S; e += TPB) {
            const int sl = (e / S + 1) << (g + 1), r = e % S;
            const i64 il = p0 + sl;
            if (il >= 0 && (il << l) < N) {
                const i64 m = il >> g, j = m / 2;
                const Cx<double> *al = ag.f[g] + j * 2 * ss + r;
                Cx<double> acc[NR];
#pragma unroll
                for (int q = 0; q < NR; q++) acc[q] = lb[q][sl][r];
                if (m >= 1) {
#pragma unroll
                    for (int c = 0; c < S; c++) {
                        const Cx<double> a = al[(i64)c * S];
#pragma unroll
                        for (int q = 0; q < NR; q++) cfms(acc[q], a, lb[q][sl - h][c]);
                    }
                }
                if (m + 1 < M) {
#pragma unroll
                    for (int c = 0; c < S; c++) {
                        const Cx<double> gm = al[ss + (i64)c * S];
#pragma unroll
                        for (int q = 0; q < NR; q++) cfms(acc[q], gm, lb[q][sl + h][c]);
                    }
                }
#pragma unroll
                for (int q = 0; q < NR; q++) lb[q][sl][r] = acc[q];
            }
        }
        __syncthreads();
    }
    for (int e = t; e < (ROWS - (1 << G)) * S * NR; e += TPB) {
        const int q = e / ((ROWS - (1 << G)) * S), rem = e % ((ROWS - (1 << G)) * S);
        const int sl = (1 << G) + rem / S, r = rem % S;
        const i64 bi = (p0 + sl) << l;
        if (bi < N) bout[q * vs + bi * S + r] = lb[q][sl][r];
    }
}

// The grouped levels with their factor loads issued at launch (s = 4, 8).
// In k_bcr_bwd_multi and k_bcr_fwd_multi every phase loads its factor blocks
// after the previous phase's barrier, so each of the G phases waits for a
// full HBM round trip, while the top phases keep most threads idle.  Here
// phase 0 runs on all TPB threads and each phase g >= 1 on a thread range of
// its own, [TPB - 2 TPB/2^g, TPB - TPB/2^g), the same rows in the same
// per-row order.  A thread loads the factor blocks (and, backward, the
// right-hand sides) of the first phase it takes part in before the launch's
// first barrier; backward, a thread of a phase g >= 1 loads its phase-0 row
// right after solving its phase-g row, forward it holds both rows' factors
// from the start.  Results are bit-identical to the kernels above.
template <int G> __device__ __forceinline__ int bcr_phase_of(int t)
{
    int gm = 0;
#pragma unroll
    for (int g = 1; g < G; g++)
        if (t >= TPB - 2 * (TPB >> g) && t < TPB - (TPB >> g)) gm = g;
    return gm;
}
__device__ __forceinline__ const Cx<double> *bcr_level_ptr(const BcrLevels &f, int g)
{
    return g == 0 ? f.f[0] : g == 1 ? f.f[1] : g == 2 ? f.f[2] : f.f[3];
}

template <int NR, int S, int G>
__device__ __forceinline__ void k_bcr_bwd_multi_pf_b(const VB vb, i64 N, int l, BcrLevels dpq, BcrLevels bl,
                                                     Cx<double> *x,
                                                          i64 vs, const Cx<double> *__restrict__ tauc, int ntau,
                                                          i64 n1, i64 ldt, Cx<double> *__restrict__ dpart)
{
    constexpr int U = TPB / ((1 << (G - 1)) * S);
    constexpr int SLOTS = U * (1 << G) + 1;
    constexpr i64 ss = (i64)S * S;
    __shared__ Cx<double> lx[NR][SLOTS][S];
    const int t = threadIdx.x;
    const i64 u0 = (i64)vb.bx * U;
    for (int e = t; e < (U + 1) * S * NR; e += TPB) {
        const int q = e / ((U + 1) * S), rem = e % ((U + 1) * S), u = rem / S, r = rem % S;
        const i64 i = ((u0 + u) << G) << l;
        lx[q][u << G][r] = i < N ? x[q * vs + i * S + r] : cx0();
    }
    // the row this thread solves in one phase: factor columns, right-hand
    // sides, its LDS slot and output offset
    Cx<double> fd[S], fp[S], fq[S], fb[NR][S];
    int slot = 0, rl = 0;
    i64 ob = 0;
    bool ok = false, okq = false;
    // phase g's thread index tg (tg < TPB >> g = U 2^(G-1-g) S)
    auto load = [&](int g, int tg) {
        const int R = 1 << (G - 1 - g);
        const int rr = tg / S, r = tg % S, u = rr / R, k = rr % R;
        const i64 hg = (i64)1 << (l + g), M = (N + hg - 1) >> (l + g);
        const i64 m = ((u0 + u) << (G - g)) + 2 * k + 1;
        ok = m < M;
        okq = m + 1 < M;
        slot = (u << G) + ((2 * k + 1) << g);
        rl = r;
        if (ok) {
            const i64 j = (m - 1) / 2;
            ob = m * hg * S;
            const Cx<double> *D = bcr_level_ptr(dpq, g) + j * 3 * ss + r, *P = D + ss, *Qm = P + ss;
            const Cx<double> *b = bcr_level_ptr(bl, g);
#pragma unroll
            for (int c = 0; c < S; c++) {
                fd[c] = D[(i64)c * S];
                fp[c] = P[(i64)c * S];
            }
            if (okq) {
#pragma unroll
                for (int c = 0; c < S; c++) fq[c] = Qm[(i64)c * S];
            }
#pragma unroll
            for (int q = 0; q < NR; q++)
#pragma unroll
                for (int c = 0; c < S; c++) fb[q][c] = b[q * vs + ob + c];
        }
    };
    auto solve = [&](int g) {
        if (!ok) return;
        Cx<double> acc[NR];
#pragma unroll
        for (int q = 0; q < NR; q++) acc[q] = cx0();
#pragma unroll
        for (int c = 0; c < S; c++) {
#pragma unroll
            for (int q = 0; q < NR; q++) {
                cfma(acc[q], fd[c], fb[q][c]);
                cfms(acc[q], fp[c], lx[q][slot - (1 << g)][c]);
            }
        }
        if (okq) {
#pragma unroll
            for (int c = 0; c < S; c++) {
#pragma unroll
                for (int q = 0; q < NR; q++) cfms(acc[q], fq[c], lx[q][slot + (1 << g)][c]);
            }
        }
#pragma unroll
        for (int q = 0; q < NR; q++) {
            x[q * vs + ob + rl] = acc[q];
            lx[q][slot][rl] = acc[q];
        }
    };
    const int gm = bcr_phase_of<G>(t);
    if (gm > 0) load(gm, t - (TPB - 2 * (TPB >> gm)));
    else load(0, t);
    __syncthreads();
#pragma unroll
    for (int g = G - 1; g >= 1; g--) {
        if (gm == g) {
            solve(g);
            load(0, t);
        }
        __syncthreads();
    }
    solve(0);
    // the dense rows' partial dots, as in k_bcr_bwd_multi
    if (tauc) {
        __syncthreads();
        __shared__ Cx<double> dred[16][NR][TPB / 64];
        const int lane = t & 63, w = t >> 6;
        for (int tt = 0; tt < ntau; tt++) {
            Cx<double> p[NR];
#pragma unroll
            for (int q = 0; q < NR; q++) p[q] = cx0();
            for (int e = t; e < U * (1 << G) * S; e += TPB) {
                const int sl = e / S, r = e % S;
                const i64 el = ((u0 << G) + sl) * S + r;
                if (el < n1 && ((u0 << G) + sl) < N) {
                    const Cx<double> tq = tauc[el + (i64)tt * ldt];
#pragma unroll
                    for (int q = 0; q < NR; q++) {
                        const Cx<double> v = lx[q][sl][r];
                        p[q].re = ffma(tq.re, v.re, p[q].re);
                        p[q].re = ffma(tq.im, v.im, p[q].re);
                        p[q].im = ffma(tq.re, v.im, p[q].im);
                        p[q].im = ffma(-tq.im, v.re, p[q].im);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NR; q++) {
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) {
                    p[q].re += __shfl_xor(p[q].re, o, 64);
                    p[q].im += __shfl_xor(p[q].im, o, 64);
                }
                if (lane == 0) dred[tt][q][w] = p[q];
            }
        }
        __syncthreads();
        if (t < ntau * NR) {
            const int tt = t / NR, q = t % NR;
            Cx<double> s = dred[tt][q][0];
#pragma unroll
            for (int v = 1; v < TPB / 64; v++) { s.re += dred[tt][q][v].re; s.im += dred[tt][q][v].im; }
            dpart[(i64)vb.bx * ntau * NR + t] = s;
        }
    }
}
template <int NR, int S, int G>
__global__ __launch_bounds__(TPB) void k_bcr_bwd_multi_pf(i64 N, int l, BcrLevels dpq, BcrLevels bl, Cx<double> *x,
                                                          i64 vs, const Cx<double> *__restrict__ tauc, int ntau,
                                                          i64 n1, i64 ldt, Cx<double> *__restrict__ dpart)
{
    k_bcr_bwd_multi_pf_b<NR, S, G>(VB_HW, N, l, dpq, bl, x, vs, tauc, ntau, n1, ldt, dpart);
}

// f0 (one right-hand side, instead of x0): the caller's vector in complex64,
// widened exactly on load.
template <int NR, int S, int G>
__device__ __forceinline__ void k_bcr_fwd_multi_pf_b(const VB vb, i64 N, int l, i64 n1, BcrLevels ag,
                                                          const Cx<double> *__restrict__ x0,
                                                          const Cx<double> *__restrict__ x1,
                                                          const Cx<double> *__restrict__ bin, Cx<double> *bout,
                                                          i64 vs, const Cx<float> *__restrict__ f0)
{
    constexpr int U = TPB / ((1 << (G - 1)) * S) - 1;
    constexpr int ROWS = (U + 1) << G;
    constexpr i64 ss = (i64)S * S;
    __shared__ Cx<double> lb[NR][ROWS][S];
    const int t = threadIdx.x;
    const i64 p0 = ((i64)vb.bx * U - 1) << G;
    const Cx<double> *xs[2] = {x0, x1};
    for (int e = t; e < ROWS * S * NR; e += TPB) {
        const int q = e / (ROWS * S), rem = e % (ROWS * S), sl = rem / S, r = rem % S;
        const i64 il = p0 + sl, bi = il << l;
        Cx<double> v = cx0();
        if (il >= 0 && bi < N) {
            if (f0) {
                const i64 el = bi * S + r;
                if (el < n1) v = cx((double)f0[el].re, (double)f0[el].im);
            } else if (x0) {
                const i64 el = bi * S + r;
                if (el < n1) v = xs[q][el];
            } else
                v = bin[q * vs + bi * S + r];
        }
        lb[q][sl][r] = v;
    }
    // one even row of a phase: its slot, lane, both neighbours' presence and
    // the factor columns alpha (left) and gamma (right)
    struct Row {
        Cx<double> a[S], c[S];
        int sl, r;
        bool ok, left, right;
    };
    auto load = [&](Row &w, int g, int tg) {
        const int nrow = (ROWS >> (g + 1)) - 1;
        w.ok = w.left = w.right = false;
        if (tg >= nrow * S) return;
        w.sl = (tg / S + 1) << (g + 1);
        w.r = tg % S;
        const i64 il = p0 + w.sl;
        if (il < 0 || (il << l) >= N) return;
        const int lv = l + g;
        const i64 M = (N + ((i64)1 << lv) - 1) >> lv, m = il >> g, j = m / 2;
        w.ok = true;
        w.left = m >= 1;
        w.right = m + 1 < M;
        const Cx<double> *al = bcr_level_ptr(ag, g) + j * 2 * ss + w.r;
        if (w.left) {
#pragma unroll
            for (int c = 0; c < S; c++) w.a[c] = al[(i64)c * S];
        }
        if (w.right) {
#pragma unroll
            for (int c = 0; c < S; c++) w.c[c] = al[ss + (i64)c * S];
        }
    };
    auto solve = [&](const Row &w, int g) {
        if (!w.ok) return;
        const int h = 1 << g;
        Cx<double> acc[NR];
#pragma unroll
        for (int q = 0; q < NR; q++) acc[q] = lb[q][w.sl][w.r];
        if (w.left) {
#pragma unroll
            for (int c = 0; c < S; c++) {
#pragma unroll
                for (int q = 0; q < NR; q++) cfms(acc[q], w.a[c], lb[q][w.sl - h][c]);
            }
        }
        if (w.right) {
#pragma unroll
            for (int c = 0; c < S; c++) {
#pragma unroll
                for (int q = 0; q < NR; q++) cfms(acc[q], w.c[c], lb[q][w.sl + h][c]);
            }
        }
#pragma unroll
        for (int q = 0; q < NR; q++) lb[q][w.sl][w.r] = acc[q];
    };
    const int gm = bcr_phase_of<G>(t);
    Row w0, w1;
    load(w0, 0, t);
    w1.ok = false;
    if (gm > 0) load(w1, gm, t - (TPB - 2 * (TPB >> gm)));
    __syncthreads();
    solve(w0, 0);
    __syncthreads();
#pragma unroll
    for (int g = 1; g < G; g++) {
        if (gm == g) solve(w1, g);
        __syncthreads();
    }
    for (int e = t; e < (ROWS - (1 << G)) * S * NR; e += TPB) {
        const int q = e / ((ROWS - (1 << G)) * S), rem = e % ((ROWS - (1 << G)) * S);
        const int sl = (1 << G) + rem / S, r = rem % S;
        const i64 bi = (p0 + sl) << l;
        if (bi < N) bout[q * vs + bi * S + r] = lb[q][sl][r];
    }
}
template <int NR, int S, int G>
__global__ __launch_bounds__(TPB) void k_bcr_fwd_multi_pf(i64 N, int l, i64 n1, BcrLevels ag,
                                                          const Cx<double> *__restrict__ x0,
                                                          const Cx<double> *__restrict__ x1,
                                                          const Cx<double> *__restrict__ bin, Cx<double> *bout,
                                                          i64 vs, const Cx<float> *__restrict__ f0)
{
    k_bcr_fwd_multi_pf_b<NR, S, G>(VB_HW, N, l, n1, ag, x0, x1, bin, bout, vs, f0);
}

// The small levels l0..L-1, the final s x s solve and the backward levels
// L-1..l0 in one workgroup of 1024 threads (every level here has at most
// 1024 rows of work), separated by workgroup barriers.  off[2l] / off[2l+1]:
// block offsets of level l's even / odd factor blocks, off[2L]: final block
// (read once into LDS).
template <int NR, int S>
__global__ __launch_bounds__(1024) void k_bcr_tail(int s, i64 N, int l0, int L, const i64 *__restrict__ off,
                                                   const Cx<double> *__restrict__ F, Cx<double> *b, Cx<double> *x,
                                                   i64 vs)
{
    __shared__ i64 loff[2 * 64 + 1];
    const int t = threadIdx.x;
    const int sv = S > 0 ? S : s;
    const i64 ss = (i64)sv * sv;
    if (t <= 2 * L) loff[t] = off[t];
    __syncthreads();
    for (int l = l0; l < L; l++) {
        const i64 h = (i64)1 << l, M = (N + h - 1) >> l, ne = (M + 1) / 2;
        const Cx<double> *AG = F + loff[2 * l] * ss;
        for (i64 g = t; g < ne * sv; g += 1024) bcr_fwd_row<NR, S>(s, M, h, g / sv, (int)(g % sv), AG, b, vs);
        __syncthreads();
    }
    if (t < sv) {
        const Cx<double> *D = F + loff[2 * L] * ss + t;
        Cx<double> acc[NR];
#pragma unroll
        for (int q = 0; q < NR; q++) acc[q] = cx0();
#pragma unroll
        for (int c = 0; c < sv; c++) {
            const Cx<double> dd = D[(i64)c * sv];
#pragma unroll
            for (int q = 0; q < NR; q++) cfma(acc[q], dd, b[q * vs + c]);
        }
#pragma unroll
        for (int q = 0; q < NR; q++) x[q * vs + t] = acc[q];
    }
    __syncthreads();
    for (int l = L - 1; l >= l0; l--) {
        const i64 h = (i64)1 << l, M = (N + h - 1) >> l, no = M / 2;
        const Cx<double> *DPQ = F + loff[2 * l + 1] * ss;
        for (i64 g = t; g < no * sv; g += 1024) bcr_bwd_row<NR, S>(s, M, h, g / sv, (int)(g % sv), DPQ, b, x, vs);
        __syncthreads();
    }
}

// The same tail levels with the right-hand sides resident in LDS: the rows
// the tail touches (block rows at multiples of 2^l0, M0 of them) are gathered
// once, every level reads and writes LDS, and the solution is scattered back
// at the end; x overwrites b in place (a backward row reads its own b and
// its neighbours' x only).  A thread's factor loads for a level are issued
// before the barrier that closes the previous level, so the global latency
// overlaps it.  M0 s NR <= 4096 complex values (64 KB): every level here
// has at most 1024 rows of work, ((M0 + 1) / 2) s <= 1024.
template <int NR, int S>
__global__ __launch_bounds__(1024) void k_bcr_tail_lds(i64 N, int l0, int L, const i64 *__restrict__ off,
                                                       const Cx<double> *__restrict__ F, const Cx<double> *b,
                                                       Cx<double> *x, i64 vs)
{
    extern __shared__ char smem_raw[];
    Cx<double> *lb = reinterpret_cast<Cx<double> *>(smem_raw);
    __shared__ i64 loff[2 * 64 + 1];
    constexpr int s = S;
    constexpr i64 ss = (i64)S * S;
    const int t = threadIdx.x;
    const i64 stride0 = (i64)1 << l0;
    const i64 M0 = (N + stride0 - 1) >> l0;
    if (t <= 2 * L) loff[t] = off[t];
    for (i64 e = t; e < M0 * s * NR; e += 1024) {
        const int q = (int)(e / (M0 * s));
        const i64 rem = e % (M0 * s), c = rem / s;
        const int r = (int)(rem % s);
        lb[e] = b[q * vs + c * stride0 * s + r];
    }
    auto L_ = [&](int q, i64 c, int r) -> Cx<double> & { return lb[((i64)q * M0 + c) * s + r]; };
    __syncthreads();       /* loff and the gathered rows visible */
    for (int l = l0; l < L; l++) {
        const i64 h = (i64)1 << l, M = (N + h - 1) >> l, ne = (M + 1) / 2, hc = (i64)1 << (l - l0);
        const i64 g = t;
        const bool act = g < ne * s;
        const i64 j = act ? g / s : 0;
        const int r = (int)(g % s);
        const i64 m = 2 * j;
        Cx<double> al[S], ga[S];
        if (act) {
            const Cx<double> *AG = F + loff[2 * l] * ss + j * 2 * ss + r;
#pragma unroll
            for (int c = 0; c < S; c++) { al[c] = AG[(i64)c * s]; ga[c] = AG[ss + (i64)c * s]; }
        }
        __syncthreads();   /* the previous level's writes are done */
        // barriers stay outside the divergent branches: a wave whose lanes
        // split would otherwise execute each branch's barrier
        Cx<double> acc[NR];
        if (act) {
#pragma unroll
            for (int q = 0; q < NR; q++) acc[q] = L_(q, m * hc, r);
            if (m >= 1) {
#pragma unroll
                for (int c = 0; c < S; c++)
#pragma unroll
                    for (int q = 0; q < NR; q++) cfms(acc[q], al[c], L_(q, (m - 1) * hc, c));
            }
            if (m + 1 < M) {
#pragma unroll
                for (int c = 0; c < S; c++)
#pragma unroll
                    for (int q = 0; q < NR; q++) cfms(acc[q], ga[c], L_(q, (m + 1) * hc, c));
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int q = 0; q < NR; q++) L_(q, m * hc, r) = acc[q];
        }
    }
    __syncthreads();
    {
        Cx<double> dd[S];
        if (t < s) {
            const Cx<double> *D = F + loff[2 * L] * ss + t;
#pragma unroll
            for (int c = 0; c < S; c++) dd[c] = D[(i64)c * s];
        }
        Cx<double> acc[NR];
        if (t < s) {
#pragma unroll
            for (int q = 0; q < NR; q++) acc[q] = cx0();
#pragma unroll
            for (int c = 0; c < S; c++)
#pragma unroll
                for (int q = 0; q < NR; q++) cfma(acc[q], dd[c], L_(q, 0, c));
        }
        __syncthreads();
        if (t < s) {
#pragma unroll
            for (int q = 0; q < NR; q++) L_(q, 0, t) = acc[q];
        }
    }
    for (int l = L - 1; l >= l0; l--) {
        const i64 h = (i64)1 << l, M = (N + h - 1) >> l, no = M / 2, hc = (i64)1 << (l - l0);
        const i64 g = t;
        const bool act = g < no * s;
        const i64 j = act ? g / s : 0;
        const int r = (int)(g % s);
        const i64 m = 2 * j + 1;
        // D^-1 and P are loaded ahead of the barrier; Q after it (registers:
        // three S-blocks of factors across a barrier exceed the 128 VGPRs of
        // a 1024-thread block)
        Cx<double> dv[S], pv[S];
        const Cx<double> *D = F + loff[2 * l + 1] * ss + j * 3 * ss + r;
        if (act) {
#pragma unroll
            for (int c = 0; c < S; c++) { dv[c] = D[(i64)c * s]; pv[c] = D[ss + (i64)c * s]; }
        }
        __syncthreads();
        Cx<double> acc[NR];
        if (act) {
#pragma unroll
            for (int q = 0; q < NR; q++) acc[q] = cx0();
#pragma unroll
            for (int c = 0; c < S; c++)
#pragma unroll
                for (int q = 0; q < NR; q++) {
                    cfma(acc[q], dv[c], L_(q, m * hc, c));
                    cfms(acc[q], pv[c], L_(q, (m - 1) * hc, c));
                }
            if (m + 1 < M) {
#pragma unroll
                for (int c = 0; c < S; c++) {
                    const Cx<double> qq = D[2 * ss + (i64)c * s];
#pragma unroll
                    for (int q = 0; q < NR; q++) cfms(acc[q], qq, L_(q, (m + 1) * hc, c));
                }
            }
        }
        __syncthreads();
        if (act) {
#pragma unroll
            for (int q = 0; q < NR; q++) L_(q, m * hc, r) = acc[q];
        }
    }
    __syncthreads();
    for (i64 e = t; e < M0 * s * NR; e += 1024) {
        const int q = (int)(e / (M0 * s));
        const i64 rem = e % (M0 * s), c = rem / s;
        const int r = (int)(rem % s);
        x[q * vs + c * stride0 * s + r] = lb[e];
    }
}

// k_bcr_tail_lds with every step on threads of its own and its factor
// blocks loaded at launch: step k < F = L - l0 is forward level l0 + k
// ((M_l + 1) / 2 even rows), step F the final block (one row), step F + 1 + k
// backward level L - 1 - k (M_l / 2 odd rows); a step's rows take S
// consecutive threads each, the steps one range after another.  So the
// chain is one HBM round trip (the gather and every factor block) and then
// LDS work between barriers, where k_bcr_tail_lds waits for each level's
// factors in turn.  Every row's arithmetic is k_bcr_tail_lds's, in its
// order.  The host launches it with 256 threads when the steps' rows fit
// (S (sum of M_l + 1) <= 256).
template <int NR, int S>
__device__ __forceinline__ void k_bcr_tail_pf_b(const VB vb, i64 N, int l0, int L, const i64 *__restrict__ off,
                                                     const Cx<double> *__restrict__ F,
                                                     const Cx<double> *__restrict__ b, Cx<double> *x, i64 vs)
{
    extern __shared__ char smem_raw[];
    Cx<double> *lb = reinterpret_cast<Cx<double> *>(smem_raw);
    constexpr i64 ss = (i64)S * S;
    const int t = threadIdx.x;
    const i64 stride0 = (i64)1 << l0;
    const i64 M0 = (N + stride0 - 1) >> l0;
    const int nf = L - l0, nsteps = 2 * nf + 1;
    for (i64 e = t; e < M0 * S * NR; e += 256) {
        const int q = (int)(e / (M0 * S));
        const i64 rem = e % (M0 * S), c = rem / S;
        const int r = (int)(rem % S);
        lb[e] = b[q * vs + c * stride0 * S + r];
    }
    auto L_ = [&](int q, i64 c, int r) -> Cx<double> & { return lb[((i64)q * M0 + c) * S + r]; };
    // this thread's step, row and lane
    int step = -1, r = 0;
    i64 j = 0;
    {
        int base = 0;
        for (int k = 0; k < nsteps && step < 0; k++) {
            i64 rows;
            if (k < nf) {
                const int l = l0 + k;
                rows = (((N + ((i64)1 << l) - 1) >> l) + 1) / 2;
            } else if (k == nf)
                rows = 1;
            else {
                const int l = L - 1 - (k - nf - 1);
                rows = ((N + ((i64)1 << l) - 1) >> l) / 2;
            }
            if (t < base + rows * S) {
                step = k;
                j = (t - base) / S;
                r = (t - base) % S;
            }
            base += (int)(rows * S);
        }
    }
    // its factor columns: forward alpha, gamma; final D; backward Dinv, P, Q
    Cx<double> f0[S], f1[S], f2[S];
    int lv = 0;
    i64 M = 1, m = 0;
    if (step >= 0 && step < nf) {
        lv = l0 + step;
        M = (N + ((i64)1 << lv) - 1) >> lv;
        m = 2 * j;
        const Cx<double> *AG = F + off[2 * lv] * ss + j * 2 * ss + r;
#pragma unroll
        for (int c = 0; c < S; c++) { f0[c] = AG[(i64)c * S]; f1[c] = AG[ss + (i64)c * S]; }
    } else if (step == nf) {
        const Cx<double> *D = F + off[2 * L] * ss + r;
#pragma unroll
        for (int c = 0; c < S; c++) f0[c] = D[(i64)c * S];
    } else if (step > nf) {
        lv = L - 1 - (step - nf - 1);
        M = (N + ((i64)1 << lv) - 1) >> lv;
        m = 2 * j + 1;
        const Cx<double> *D = F + off[2 * lv + 1] * ss + j * 3 * ss + r;
#pragma unroll
        for (int c = 0; c < S; c++) { f0[c] = D[(i64)c * S]; f1[c] = D[ss + (i64)c * S]; }
        if (m + 1 < M) {
#pragma unroll
            for (int c = 0; c < S; c++) f2[c] = D[2 * ss + (i64)c * S];
        }
    }
    const i64 hc = (i64)1 << (lv - l0);
    __syncthreads();   /* the gathered rows visible */
    // barriers stay outside the divergent branches
    for (int k = 0; k < nsteps; k++) {
        Cx<double> acc[NR];
        if (step == k) {
            if (k < nf) {
#pragma unroll
                for (int q = 0; q < NR; q++) acc[q] = L_(q, m * hc, r);
                if (m >= 1) {
#pragma unroll
                    for (int c = 0; c < S; c++)
#pragma unroll
                        for (int q = 0; q < NR; q++) cfms(acc[q], f0[c], L_(q, (m - 1) * hc, c));
                }
                if (m + 1 < M) {
#pragma unroll
                    for (int c = 0; c < S; c++)
#pragma unroll
                        for (int q = 0; q < NR; q++) cfms(acc[q], f1[c], L_(q, (m + 1) * hc, c));
                }
            } else if (k == nf) {
#pragma unroll
                for (int q = 0; q < NR; q++) acc[q] = cx0();
#pragma unroll
                for (int c = 0; c < S; c++)
#pragma unroll
                    for (int q = 0; q < NR; q++) cfma(acc[q], f0[c], L_(q, 0, c));
            } else {
#pragma unroll
                for (int q = 0; q < NR; q++) acc[q] = cx0();
#pragma unroll
                for (int c = 0; c < S; c++)
#pragma unroll
                    for (int q = 0; q < NR; q++) {
                        cfma(acc[q], f0[c], L_(q, m * hc, c));
                        cfms(acc[q], f1[c], L_(q, (m - 1) * hc, c));
                    }
                if (m + 1 < M) {
#pragma unroll
                    for (int c = 0; c < S; c++)
#pragma unroll
                        for (int q = 0; q < NR; q++) cfms(acc[q], f2[c], L_(q, (m + 1) * hc, c));
                }
            }
        }
        __syncthreads();
        if (step == k) {
            const i64 row = k == nf ? 0 : m * hc;
#pragma unroll
            for (int q = 0; q < NR; q++) L_(q, row, r) = acc[q];
        }
        __syncthreads();
    }
    for (i64 e = t; e < M0 * S * NR; e += 256) {
        const int q = (int)(e / (M0 * S));
        const i64 rem = e % (M0 * S), c = rem / S;
        const int rr = (int)(rem % S);
        x[q * vs + c * stride0 * S + rr] = lb[e];
    }
}
template <int NR, int S>
__global__ __launch_bounds__(256) void k_bcr_tail_pf(i64 N, int l0, int L, const i64 *__restrict__ off,
                                                     const Cx<double> *__restrict__ F,
                                                     const Cx<double> *__restrict__ b, Cx<double> *x, i64 vs)
{
    k_bcr_tail_pf_b<NR, S>(VB_HW, N, l0, L, off, F, b, x, vs);
}

// border, vector q: t = b2_q - dots_q; y2_q = Sinv t -> y2out + q ntau and the
// tail of y_q.  One workgroup, NR * ntau threads used.  With dpart the dots
// are first summed from nblk per-block partials (layout [block][a][q]), wave
// o handling outputs o, o + waves, ... in a fixed order.
// b2f / ytf (one right-hand side): the dense rows' right-hand side read and
// the tail of y written in complex64 instead of b2_0 / ytail0.
template <int NR>
__device__ __forceinline__ void k_border_small_b(const VB vb, int ntau, const Cx<double> *b2_0,
                                                 const Cx<double> *b2_1, const Cx<double> *dots0,
                               const Cx<double> *dots1, const Cx<double> *__restrict__ Sinv, Cx<double> *y2out,
                               Cx<double> *ytail0, Cx<double> *ytail1, const Cx<double> *__restrict__ dpart,
                               i64 nblk, const Cx<float> *b2f, Cx<float> *ytf)
{
    extern __shared__ char sm[];
    Cx<double> *tt = reinterpret_cast<Cx<double> *>(sm);
    __shared__ Cx<double> dd[2 * 16];
    const int a = threadIdx.x % ntau, q = threadIdx.x / ntau;
    const bool act = q < NR;
    if (dpart) {
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
        for (int o = wv; o < ntau * NR; o += nw) {
            Cx<double> s = cx0();
            for (i64 b = lane; b < nblk; b += 64) {
                const Cx<double> v = dpart[b * ntau * NR + o];
                s.re += v.re; s.im += v.im;
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) {
                s.re += __shfl_xor(s.re, off, 64);
                s.im += __shfl_xor(s.im, off, 64);
            }
            if (lane == 0) dd[o] = s;       // o = a' NR + q'
        }
        __syncthreads();
    }
    if (act) {
        const Cx<double> b2 = b2f ? cx((double)b2f[a].re, (double)b2f[a].im) : (q == 0 ? b2_0 : b2_1)[a];
        const Cx<double> d = dpart ? dd[a * NR + q] : (q == 0 ? dots0 : dots1)[a];
        tt[q * ntau + a] = cx(b2.re - d.re, b2.im - d.im);
    }
    __syncthreads();
    if (act) {
        Cx<double> acc = cx0();
        for (int c = 0; c < ntau; c++) cfma(acc, Sinv[(i64)c * ntau + a], tt[q * ntau + c]);
        y2out[q * ntau + a] = acc;
        if (ytf) ytf[a] = Cx<float>{(float)acc.re, (float)acc.im};
        else (q == 0 ? ytail0 : ytail1)[a] = acc;
    }
}
template <int NR>
__global__ void k_border_small(int ntau, const Cx<double> *b2_0, const Cx<double> *b2_1, const Cx<double> *dots0,
                               const Cx<double> *dots1, const Cx<double> *__restrict__ Sinv, Cx<double> *y2out,
                               Cx<double> *ytail0, Cx<double> *ytail1, const Cx<double> *__restrict__ dpart,
                               i64 nblk, const Cx<float> *b2f, Cx<float> *ytf)
{
    k_border_small_b<NR>(VB_HW, ntau, b2_0, b2_1, dots0, dots1, Sinv, y2out, ytail0, ytail1, dpart, nblk, b2f, ytf);
}

// y_q[r] = xb_q[r] - sum_a Y[a n1 + r] y2_q[a], r < n1; with y0f (one
// right-hand side) y_0 is written in complex64 there instead.
template <int NR>
__device__ __forceinline__ void k_border_update_b(const VB vb, i64 n1, int ntau, const Cx<double> *__restrict__ xb,
                                                  i64 vs,
                                                       const Cx<double> *__restrict__ Y,
                                                       const Cx<double> *__restrict__ y2, Cx<double> *y0,
                                                       Cx<double> *y1, Cx<float> *y0f)
{
    for (i64 i = (i64)vb.bx * TPB + threadIdx.x; i < n1; i += (i64)vb.gx * TPB) {
        Cx<double> acc[NR];
#pragma unroll
        for (int q = 0; q < NR; q++) acc[q] = xb[q * vs + i];
        for (int a = 0; a < ntau; a++) {
            const Cx<double> yv = Y[(i64)a * n1 + i];
#pragma unroll
            for (int q = 0; q < NR; q++) cfms(acc[q], yv, y2[q * ntau + a]);
        }
        if (y0f) y0f[i] = Cx<float>{(float)acc[0].re, (float)acc[0].im};
        else y0[i] = acc[0];
        if (NR == 2) y1[i] = acc[NR - 1];
    }
}
template <int NR>
__global__ __launch_bounds__(TPB) void k_border_update(i64 n1, int ntau, const Cx<double> *__restrict__ xb, i64 vs,
                                                       const Cx<double> *__restrict__ Y,
                                                       const Cx<double> *__restrict__ y2, Cx<double> *y0,
                                                       Cx<double> *y1, Cx<float> *y0f)
{
    k_border_update_b<NR>(VB_HW, n1, ntau, xb, vs, Y, y2, y0, y1, y0f);
}

static inline int blocks_for(i64 n) { i64 b = (n + TPB - 1) / TPB; return (int)(b < 1 ? 1 : b); }
static inline int gs_blocks(i64 n) { i64 b = (n + TPB - 1) / TPB; if (b > 8192) b = 8192; return (int)(b < 1 ? 1 : b); }
static inline int chk(void) { return hipGetLastError() == hipSuccess ? 0 : -1; }

extern "C" {

int bbjd_o_band_mv(hipStream_t s, int64_t n1, int64_t nb, int bs, int hw, const void *band, const void *x, void *y,
                   int single)
{
    if (n1 <= 0) return 0;
    if (single)
        bbjd::gl<&k_band_mv_b<float>, TPB>(k_band_mv<float>, dim3(blocks_for(n1)), 0, s, (i64)n1, (i64)nb, bs, hw,
                           (const Cx<float> *)band, (const Cx<float> *)x, (Cx<float> *)y);
    else
        bbjd::gl<&k_band_mv_b<double>, TPB>(k_band_mv<double>, dim3(blocks_for(n1)), 0, s, (i64)n1, (i64)nb, bs, hw,
                           (const Cx<double> *)band, (const Cx<double> *)x, (Cx<double> *)y);
    return chk();
}

int bbjd_o_band_ab(hipStream_t s, int64_t n, int64_t n1, int64_t nb, int bs, int hw, const void *band,
                   const void *dg, int ntau, const void *tauc, const void *x, void *ya, void *yb, void *part,
                   int64_t part_cap, int single)
{
    return bbjd_o_band_abr(s, n, n1, nb, bs, hw, band, dg, ntau, tauc, x, ya, yb, part, part_cap, single, 0, 0, 0,
                           0, nullptr, nullptr, 0, nullptr);
}

int bbjd_o_band_abr(hipStream_t s, int64_t n, int64_t n1, int64_t nb, int bs, int hw, const void *band,
                    const void *dg, int ntau, const void *tauc, const void *x, void *ya, void *yb, void *part,
                    int64_t part_cap, int single, double are, double aim, double bre, double bim, void *r,
                    void *rpart, int64_t rpart_cap, double *rnorm)
{
    if (n <= 0) return 0;
    const i64 nblk = blocks_for(n);
    const int res = r != nullptr;
    if (ntau < 0 || ntau > AB_MAX_TAU || (ntau > 0 && nblk * ntau > part_cap)) return -3;
    if (res && (single || ntau >= AB_MAX_TAU || !rpart || !rnorm || nblk > rpart_cap)) return -3;
    const int fin_threads = 64 * (ntau + res);
#define BBJD_AB_K(R, B, H)                                                                                      \
    bbjd::gl<&k_band_ab_b<R, B, H>, TPB>(k_band_ab<R, B, H>, dim3(nblk), 0, s, (i64)n, (i64)n1, (i64)nb, bs, hw,       \
                       (const Cx<R> *)band, (const Cx<R> *)dg, ntau, (const Cx<R> *)tauc, (const Cx<R> *)x,       \
                       (Cx<R> *)ya, (Cx<R> *)yb, (Cx<R> *)part, Cx<R>{(R)are, (R)aim}, Cx<R>{(R)bre, (R)bim},     \
                       (Cx<R> *)r, (R *)rpart)
/* the dense rows' finish, groupable for the block sizes of up to 3 dense
 * rows with the residual (one wave per row) */
#define BBJD_TAU_FIN(R, BSZ)                                                                                    \
    bbjd::gl<&k_tau_finish_b<R>, BSZ>(k_tau_finish<R>, dim3(1), 0, s, (i64)nblk, ntau, (i64)n1,                 \
                                      (const Cx<R> *)part, (Cx<R> *)ya, (const Cx<R> *)yb,                      \
                                      Cx<R>{(R)are, (R)aim}, Cx<R>{(R)bre, (R)bim}, (Cx<R> *)r,                 \
                                      (const R *)rpart, rnorm)
#define BBJD_AB(R)                                                                                              \
    if (bs == 8 && hw == 1) BBJD_AB_K(R, 8, 1);                                                                 \
    else if (bs == 4 && hw == 1) BBJD_AB_K(R, 4, 1);                                                            \
    else if (bs == 8 && hw == 2) BBJD_AB_K(R, 8, 2);                                                            \
    else BBJD_AB_K(R, 0, 0);                                                                                    \
    if (fin_threads > 0)                                                                                        \
        switch (fin_threads) {                                                                                  \
        case 64: BBJD_TAU_FIN(R, 64); break;                                                                    \
        case 128: BBJD_TAU_FIN(R, 128); break;                                                                  \
        case 192: BBJD_TAU_FIN(R, 192); break;                                                                  \
        case 256: BBJD_TAU_FIN(R, 256); break;                                                                  \
        default:                                                                                                \
            BBJD_SL(s, (k_tau_finish<R>), dim3(1), dim3(fin_threads), 0, (i64)nblk, ntau, (i64)n1,              \
                    (const Cx<R> *)part, (Cx<R> *)ya, (const Cx<R> *)yb, Cx<R>{(R)are, (R)aim},                 \
                    Cx<R>{(R)bre, (R)bim}, (Cx<R> *)r, (const R *)rpart, rnorm);                                \
        }
    if (single) { BBJD_AB(float) } else { BBJD_AB(double) }
#undef BBJD_AB
#undef BBJD_AB_K
#undef BBJD_TAU_FIN
    return chk();
}

int bbjd_o_diag_mv(hipStream_t s, int64_t n, const void *dg, const void *x, void *y, int single)
{
    if (single)
        bbjd::gl<&k_diag_mv_b<float>, TPB>(k_diag_mv<float>, dim3(gs_blocks(n)), 0, s, (i64)n, (const Cx<float> *)dg,
                           (const Cx<float> *)x, (Cx<float> *)y);
    else
        bbjd::gl<&k_diag_mv_b<double>, TPB>(k_diag_mv<double>, dim3(gs_blocks(n)), 0, s, (i64)n, (const Cx<double> *)dg,
                           (const Cx<double> *)x, (Cx<double> *)y);
    return chk();
}

int bbjd_o_bj_apply(hipStream_t s, int64_t n, int bs, const void *inv, const void *x, void *y, int single)
{
    if (single)
        bbjd::gl<&k_bj_apply_b<float>, TPB>(k_bj_apply<float>, dim3(blocks_for(n)), 0, s, (i64)n, bs,
                           (const Cx<float> *)inv, (const Cx<float> *)x, (Cx<float> *)y);
    else
        bbjd::gl<&k_bj_apply_b<double>, TPB>(k_bj_apply<double>, dim3(blocks_for(n)), 0, s, (i64)n, bs,
                           (const Cx<double> *)inv, (const Cx<double> *)x, (Cx<double> *)y);
    return chk();
}

// Instantiation by (right-hand sides, super-block size): s = 4, 8, 16 are
// compiled with fixed loops, other sizes take the generic S = 0 kernels.
#define BBJD_BCR_DISPATCH(KERNEL_CALL)                      \
    if (nrhs == 2) {                                        \
        switch (s) {                                        \
        case 4: KERNEL_CALL(2, 4); break;                   \
        case 8: KERNEL_CALL(2, 8); break;                   \
        case 16: KERNEL_CALL(2, 16); break;                 \
        default: KERNEL_CALL(2, 0); break;                  \
        }                                                   \
    } else {                                                \
        switch (s) {                                        \
        case 4: KERNEL_CALL(1, 4); break;                   \
        case 8: KERNEL_CALL(1, 8); break;                   \
        case 16: KERNEL_CALL(1, 16); break;                 \
        default: KERNEL_CALL(1, 0); break;                  \
        }                                                   \
    }

int bbjd_o_bcr_fwd(hipStream_t st, int s, int64_t M, int64_t h, int64_t ne, const void *AG, void *b, int nrhs,
                   int64_t vs)
{
    if (ne <= 0) return 0;
#define BBJD_FWD(NR, S)                                                                                         \
    BBJD_SL(st, (k_bcr_fwd<NR, S>), dim3(blocks_for(ne * s)), dim3(TPB), 0, s, (i64)M, (i64)h, (i64)ne, \
                       (const Cx<double> *)AG, (Cx<double> *)b, (i64)vs)
    BBJD_BCR_DISPATCH(BBJD_FWD)
#undef BBJD_FWD
    return chk();
}

int bbjd_o_bcr_fwd0(hipStream_t st, int s, int64_t N, int64_t n1, const void *AG, const void *x0, const void *x1,
                    void *b, int nrhs, int64_t vs)
{
    if (N <= 0) return 0;
    if (s > 64 || (nrhs == 2 && !x1)) return -2;
#define BBJD_FWD0(NR, S)                                                                                        \
    BBJD_SL(st, (k_bcr_fwd0<NR, S>), dim3(blocks_for(N * s)), dim3(TPB), 0, s, (i64)N, (i64)n1,        \
                       (const Cx<double> *)AG, (const Cx<double> *)x0, (const Cx<double> *)(x1 ? x1 : x0),       \
                       (Cx<double> *)b, (i64)vs)
    BBJD_BCR_DISPATCH(BBJD_FWD0)
#undef BBJD_FWD0
    return chk();
}

int bbjd_o_bcr_bwd(hipStream_t st, int s, int64_t M, int64_t h, int64_t no, const void *DPQ, const void *b, void *x,
                   int nrhs, int64_t vs)
{
    if (no <= 0) return 0;
#define BBJD_BWD(NR, S)                                                                                         \
    BBJD_SL(st, (k_bcr_bwd<NR, S>), dim3(blocks_for(no * s)), dim3(TPB), 0, s, (i64)M, (i64)h, (i64)no, \
                       (const Cx<double> *)DPQ, (const Cx<double> *)b, (Cx<double> *)x, (i64)vs)
    BBJD_BCR_DISPATCH(BBJD_BWD)
#undef BBJD_BWD
    return chk();
}

// Fused level groups: s in {4, 8, 16}, G in {2, 3, 4}; -3 for other shapes
// (the caller then launches level by level).
#define BBJD_BCR_GROUP_DISPATCH(CALL)                                                    \
    switch (s * 16 + G * 2 + (nrhs == 2)) {                                              \
    case 4 * 16 + 4: CALL(1, 4, 2); break;   case 4 * 16 + 5: CALL(2, 4, 2); break;      \
    case 4 * 16 + 6: CALL(1, 4, 3); break;   case 4 * 16 + 7: CALL(2, 4, 3); break;      \
    case 4 * 16 + 8: CALL(1, 4, 4); break;   case 4 * 16 + 9: CALL(2, 4, 4); break;      \
    case 8 * 16 + 4: CALL(1, 8, 2); break;   case 8 * 16 + 5: CALL(2, 8, 2); break;      \
    case 8 * 16 + 6: CALL(1, 8, 3); break;   case 8 * 16 + 7: CALL(2, 8, 3); break;      \
    case 8 * 16 + 8: CALL(1, 8, 4); break;   case 8 * 16 + 9: CALL(2, 8, 4); break;      \
    case 16 * 16 + 4: CALL(1, 16, 2); break; case 16 * 16 + 5: CALL(2, 16, 2); break;    \
    case 16 * 16 + 6: CALL(1, 16, 3); break; case 16 * 16 + 7: CALL(2, 16, 3); break;    \
    case 16 * 16 + 8: CALL(1, 16, 4); break; case 16 * 16 + 9: CALL(2, 16, 4); break;    \
    default: return -3;                                                                  \
    }

static inline bool bcr_group_ok(int s, int G) { return (s == 4 || s == 8 || s == 16) && G >= 2 && G <= 4; }

// s = 4, 8: the grouped levels with their factor loads issued at launch
// (k_bcr_*_multi_pf); 0 selects the phase-by-phase kernels (for tests)
static int g_bcr_prefetch = 1;
extern "C" void bbjd_set_bcr_prefetch(int on) { g_bcr_prefetch = on != 0; }

int bbjd_o_bcr_bwd_multi(hipStream_t st, int s, int64_t N, int l, int G, const void *const *dpq,
                         const void *const *bl, void *x, int nrhs, int64_t vs, const void *tauc, int ntau,
                         int64_t n1, int64_t ldt, void *dpart, int64_t dpart_cap, int64_t *nblk_out)
{
    if (!bcr_group_ok(s, G) || (nrhs != 1 && nrhs != 2)) return -3;
    if (nblk_out) *nblk_out = 0;
    BcrLevels f, bb;
    for (int g = 0; g < 4; g++) {
        f.f[g] = (const Cx<double> *)(g < G ? dpq[g] : nullptr);
        bb.f[g] = (const Cx<double> *)(g < G ? bl[g] : nullptr);
    }
    const i64 H = (i64)1 << (l + G), units = (N + H - 1) / H;
    // the dense rows' partial dots: only from the group that ends at level 0
    const bool dots = tauc && l == 0 && ntau > 0 && ntau <= 16 && dpart;
#define BBJD_BWDM(NR, S, GG)                                                                                   \
    {                                                                                                          \
        constexpr int U = TPB / ((1 << (GG - 1)) * S);                                                         \
        const i64 nblk = (units + U - 1) / U;                                                                  \
        const bool dd = dots && nblk * ntau * NR <= dpart_cap;                                                 \
        const Cx<double> *tc = dd ? (const Cx<double> *)tauc : nullptr;                                        \
        if constexpr (S <= 8) {                                                                                \
            if (g_bcr_prefetch) {                                                                              \
                bbjd::gl<&k_bcr_bwd_multi_pf_b<NR, S, GG>, TPB>(k_bcr_bwd_multi_pf<NR, S, GG>,                       \
                                                                dim3((unsigned)nblk), 0, st,                         \
                                   (i64)N, l, f, bb, (Cx<double> *)x, (i64)vs, tc, ntau, (i64)n1, (i64)ldt,    \
                                   (Cx<double> *)dpart);                                                       \
                if (dd && nblk_out) *nblk_out = nblk;                                                          \
                break;                                                                                         \
            }                                                                                                  \
        }                                                                                                      \
        BBJD_SL(st, (k_bcr_bwd_multi<NR, S, GG>), dim3((unsigned)nblk), dim3(TPB), 0, (i64)N, l, f, bb, \
                           (Cx<double> *)x, (i64)vs, tc, ntau, (i64)n1, (i64)ldt, (Cx<double> *)dpart);         \
        if (dd && nblk_out) *nblk_out = nblk;                                                                  \
    }
    BBJD_BCR_GROUP_DISPATCH(BBJD_BWDM)
#undef BBJD_BWDM
    return chk();
}

int bbjd_o_bcr_fwd_multi(hipStream_t st, int s, int64_t N, int64_t n1, int l, int G, const void *const *ag,
                         const void *x0, const void *x1, const void *bin, void *bout, int nrhs, int64_t vs,
                         int in_c64)
{
    if (!bcr_group_ok(s, G) || (nrhs != 1 && nrhs != 2) || (x0 && (l != 0 || (nrhs == 2 && !x1))) ||
        (!x0 && (!bin || bin == bout)))
        return -3;
    // complex64 input: one right-hand side through the prefetching kernels
    if (in_c64 && (nrhs != 1 || !x0 || s > 8 || !g_bcr_prefetch)) return -3;
    BcrLevels f;
    for (int g = 0; g < 4; g++) f.f[g] = (const Cx<double> *)(g < G ? ag[g] : nullptr);
    const i64 Ml = (N + ((i64)1 << l) - 1) >> l, units = (Ml + ((i64)1 << G) - 1) >> G;
#define BBJD_FWDM(NR, S, GG)                                                                                   \
    {                                                                                                          \
        constexpr int U = TPB / ((1 << (GG - 1)) * S) - 1;                                                     \
        const unsigned nb = (unsigned)((units + U - 1) / U);                                                   \
        const Cx<double> *a0 = (const Cx<double> *)x0, *a1 = (const Cx<double> *)(x1 ? x1 : x0);               \
        if constexpr (S <= 8) {                                                                                \
            if (g_bcr_prefetch) {                                                                              \
                const Cx<float> *f0 = in_c64 ? (const Cx<float> *)x0 : nullptr;                                \
                if (in_c64) a0 = a1 = nullptr;                                                                 \
                bbjd::gl<&k_bcr_fwd_multi_pf_b<NR, S, GG>, TPB>(k_bcr_fwd_multi_pf<NR, S, GG>, dim3(nb), 0, st,      \
                                                                (i64)N, l,                                           \
                                   (i64)n1, f, a0, a1, (const Cx<double> *)bin, (Cx<double> *)bout, (i64)vs,   \
                                   f0);                                                                        \
                break;                                                                                         \
            }                                                                                                  \
        }                                                                                                      \
        BBJD_SL(st, (k_bcr_fwd_multi<NR, S, GG>), dim3(nb), dim3(TPB), 0, (i64)N, l, (i64)n1, f, a0, a1, \
                           (const Cx<double> *)bin, (Cx<double> *)bout, (i64)vs);                              \
    }
    BBJD_BCR_GROUP_DISPATCH(BBJD_FWDM)
#undef BBJD_FWDM
    return chk();
}

// The LDS tail holds the rows of its first level in at most 60 KB (with its
// static LDS the block stays within 64 KB); it serves s = 4, 8 (s = 16 needs
// more than a 1024-thread block's registers for its prefetched factor
// blocks and keeps the global tail).
static const size_t TAIL_LDS_MAX = 60 * 1024;
static inline bool tail_lds_fits(int s, i64 N, int l, int nrhs)
{
    const i64 M = (N + ((i64)1 << l) - 1) >> l;
    return (s == 4 || s == 8) && (size_t)M * s * nrhs * sizeof(Cx<double>) <= TAIL_LDS_MAX;
}

int bbjd_o_bcr_tail_level(int s, int64_t N, int l0, int L, int nrhs)
{
    if (s != 4 && s != 8) return l0;
    for (int l = l0; l <= L; l++)
        if (tail_lds_fits(s, N, l, nrhs)) return l;
    return l0;
}

int bbjd_o_bcr_tail(hipStream_t st, int s, int64_t N, int l0, int L, const int64_t *off, const void *F, void *b,
                    void *x, int nrhs, int64_t vs)
{
    if (s > 1024 || L > 64) return -2;
    const i64 M0 = (N + ((i64)1 << l0) - 1) >> l0;
    const size_t lds = (size_t)M0 * s * nrhs * sizeof(Cx<double>);
    if (tail_lds_fits(s, N, l0, nrhs)) {
        // every step on threads of its own when they fit 256
        i64 rows = 1;
        for (int l = l0; l < L; l++) rows += (N + ((i64)1 << l) - 1) >> l;
        if (g_bcr_prefetch && rows * s <= 256) {
#define BBJD_TAIL_PF(NR, SS)                                                                                   \
    bbjd::gl<&k_bcr_tail_pf_b<NR, SS>, 256>(k_bcr_tail_pf<NR, SS>, dim3(1), lds, st, (i64)N, l0, L,                  \
                                            (const i64 *)off,                                                        \
                       (const Cx<double> *)F, (const Cx<double> *)b, (Cx<double> *)x, (i64)vs)
            if (nrhs == 2) {
                if (s == 4) BBJD_TAIL_PF(2, 4);
                else BBJD_TAIL_PF(2, 8);
            } else {
                if (s == 4) BBJD_TAIL_PF(1, 4);
                else BBJD_TAIL_PF(1, 8);
            }
#undef BBJD_TAIL_PF
            return chk();
        }
#define BBJD_TAIL_LDS(NR, SS)                                                                                  \
    BBJD_SL(st, (k_bcr_tail_lds<NR, SS>), dim3(1), dim3(1024), lds, (i64)N, l0, L, (const i64 *)off, \
                       (const Cx<double> *)F, (const Cx<double> *)b, (Cx<double> *)x, (i64)vs)
        if (nrhs == 2) {
            if (s == 4) BBJD_TAIL_LDS(2, 4);
            else BBJD_TAIL_LDS(2, 8);
        } else {
            if (s == 4) BBJD_TAIL_LDS(1, 4);
            else BBJD_TAIL_LDS(1, 8);
        }
#undef BBJD_TAIL_LDS
        return chk();
    }
#define BBJD_TAIL(NR, S)                                                                                        \
    BBJD_SL(st, (k_bcr_tail<NR, S>), dim3(1), dim3(1024), 0, s, (i64)N, l0, L, (const i64 *)off,       \
                       (const Cx<double> *)F, (Cx<double> *)b, (Cx<double> *)x, (i64)vs)
    BBJD_BCR_DISPATCH(BBJD_TAIL)
#undef BBJD_TAIL
    return chk();
}

int bbjd_o_border_small(hipStream_t st, int ntau, const void *b2_0, const void *b2_1, const void *dots0,
                        const void *dots1, const void *Sinv, void *y2, void *ytail0, void *ytail1, int nrhs,
                        const void *dpart, int64_t nblk, int c64)
{
    if (ntau <= 0 || ntau * nrhs > 1024 || (dpart && ntau > 16) || (c64 && nrhs != 1)) return -2;
    const int thr = dpart ? 1024 : ntau * nrhs <= 64 ? 64 : (ntau * nrhs <= 256 ? 256 : 1024);
    if (nrhs == 1 && (thr == 64 || thr == 256 || thr == 1024)) {
        /* groupable */
        const Cx<double> *b20 = c64 ? nullptr : (const Cx<double> *)b2_0;
        Cx<double> *yt0 = c64 ? nullptr : (Cx<double> *)ytail0;
        const Cx<float> *b2f = c64 ? (const Cx<float> *)b2_0 : nullptr;
        Cx<float> *ytf = c64 ? (Cx<float> *)ytail0 : nullptr;
#define BBJD_BORDER1(BSZ)                                                                                       \
    bbjd::gl<&k_border_small_b<1>, BSZ>(k_border_small<1>, dim3(1), (size_t)ntau * 16, st, ntau, b20, b20,     \
                                        (const Cx<double> *)dots0, (const Cx<double> *)dots0,                   \
                                        (const Cx<double> *)Sinv, (Cx<double> *)y2, yt0, yt0,                   \
                                        (const Cx<double> *)dpart, (i64)nblk, b2f, ytf)
        if (thr == 64) BBJD_BORDER1(64);
        else if (thr == 256) BBJD_BORDER1(256);
        else BBJD_BORDER1(1024);
#undef BBJD_BORDER1
    } else if (nrhs == 2)
        BBJD_SL(st, k_border_small<2>, dim3(1), dim3(thr), (size_t)2 * ntau * 16, ntau,
                           (const Cx<double> *)b2_0, (const Cx<double> *)b2_1, (const Cx<double> *)dots0,
                           (const Cx<double> *)dots1, (const Cx<double> *)Sinv, (Cx<double> *)y2,
                           (Cx<double> *)ytail0, (Cx<double> *)ytail1, (const Cx<double> *)dpart, (i64)nblk,
                           (const Cx<float> *)nullptr, (Cx<float> *)nullptr);
    else
        BBJD_SL(st, k_border_small<1>, dim3(1), dim3(thr), (size_t)ntau * 16, ntau,
                           c64 ? nullptr : (const Cx<double> *)b2_0, c64 ? nullptr : (const Cx<double> *)b2_0,
                           (const Cx<double> *)dots0, (const Cx<double> *)dots0, (const Cx<double> *)Sinv,
                           (Cx<double> *)y2, c64 ? nullptr : (Cx<double> *)ytail0,
                           c64 ? nullptr : (Cx<double> *)ytail0, (const Cx<double> *)dpart, (i64)nblk,
                           c64 ? (const Cx<float> *)b2_0 : nullptr, c64 ? (Cx<float> *)ytail0 : nullptr);
    return chk();
}

int bbjd_o_border_update(hipStream_t st, int64_t n1, int ntau, const void *xb, int64_t vs, const void *Y,
                         const void *y2, void *y0, void *y1, int nrhs, int c64)
{
    if (c64 && nrhs != 1) return -2;
    if (nrhs == 2)
        bbjd::gl<&k_border_update_b<2>, TPB>(k_border_update<2>, dim3(gs_blocks(n1)), 0, st, (i64)n1, ntau,
                           (const Cx<double> *)xb, (i64)vs, (const Cx<double> *)Y, (const Cx<double> *)y2,
                           (Cx<double> *)y0, (Cx<double> *)y1, (Cx<float> *)nullptr);
    else
        bbjd::gl<&k_border_update_b<1>, TPB>(k_border_update<1>, dim3(gs_blocks(n1)), 0, st, (i64)n1, ntau,
                           (const Cx<double> *)xb, (i64)vs, (const Cx<double> *)Y, (const Cx<double> *)y2,
                           c64 ? nullptr : (Cx<double> *)y0, c64 ? nullptr : (Cx<double> *)y0,
                           c64 ? (Cx<float> *)y0 : nullptr);
    return chk();
}

}  // extern "C"
