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

typedef long long i64;
template <typename R> struct alignas(2 * sizeof(R)) Cx { R re, im; };
template <typename R> __device__ __forceinline__ void cfma(Cx<R> &a, Cx<R> q, Cx<R> h)
{ a.re += q.re * h.re - q.im * h.im; a.im += q.re * h.im + q.im * h.re; }
template <typename R> __device__ __forceinline__ void cfms(Cx<R> &a, Cx<R> q, Cx<R> h)
{ a.re -= q.re * h.re - q.im * h.im; a.im -= q.re * h.im + q.im * h.re; }

static constexpr int TPB = 256;

template <typename R>
__global__ __launch_bounds__(TPB) void k_band_mv(i64 n1, i64 nb, int bs, int hw, const Cx<R> *__restrict__ band,
                                                 const Cx<R> *__restrict__ x, Cx<R> *__restrict__ y)
{
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
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
__global__ __launch_bounds__(TPB) void k_diag_mv(i64 n, const Cx<R> *__restrict__ dg, const Cx<R> *__restrict__ x,
                                                 Cx<R> *__restrict__ y)
{
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) {
        Cx<R> a = {0, 0};
        cfma(a, dg[i], x[i]);
        y[i] = a;
    }
}

template <typename R>
__global__ __launch_bounds__(TPB) void k_bj_apply(i64 n, int bs, const Cx<R> *__restrict__ inv,
                                                  const Cx<R> *__restrict__ x, Cx<R> *__restrict__ y)
{
    const i64 i = (i64)blockIdx.x * TPB + threadIdx.x;
    if (i >= n) return;
    const i64 bi = i / bs;
    const int r = (int)(i % bs);
    const Cx<R> *D = inv + bi * bs * bs + r;
    const Cx<R> *xs = x + bi * bs;
    Cx<R> a = {0, 0};
    for (int c = 0; c < bs; c++) cfma(a, D[(i64)c * bs], xs[c]);
    y[i] = a;
}

// forward level: evens m = 2j at stride h, j < ne.
__global__ __launch_bounds__(TPB) void k_bcr_fwd(int s, i64 M, i64 h, i64 ne, const Cx<double> *__restrict__ AG,
                                                 Cx<double> *b)
{
    const i64 gid = (i64)blockIdx.x * TPB + threadIdx.x;
    const i64 j = gid / s;
    const int r = (int)(gid % s);
    if (j >= ne) return;
    const i64 m = 2 * j, i = m * h, ss = (i64)s * s;
    Cx<double> acc = b[i * s + r];
    if (m >= 1) {
        const Cx<double> *al = AG + j * 2 * ss + r;
        const Cx<double> *bl = b + (i - h) * s;
        for (int c = 0; c < s; c++) cfms(acc, al[(i64)c * s], bl[c]);
    }
    if (m + 1 < M) {
        const Cx<double> *ga = AG + j * 2 * ss + ss + r;
        const Cx<double> *br = b + (i + h) * s;
        for (int c = 0; c < s; c++) cfms(acc, ga[(i64)c * s], br[c]);
    }
    b[i * s + r] = acc;
}

// backward level: odds m = 2j + 1 at stride h, j < no.  DPQ = (Dinv, Dinv L,
// Dinv U) per odd.  Reads b and the already-solved even x, writes x.
__global__ __launch_bounds__(TPB) void k_bcr_bwd(int s, i64 M, i64 h, i64 no, const Cx<double> *__restrict__ DPQ,
                                                 const Cx<double> *__restrict__ b, Cx<double> *x)
{
    const i64 gid = (i64)blockIdx.x * TPB + threadIdx.x;
    const i64 j = gid / s;
    const int r = (int)(gid % s);
    if (j >= no) return;
    const i64 m = 2 * j + 1, i = m * h, ss = (i64)s * s;
    const Cx<double> *D = DPQ + j * 3 * ss + r, *P = D + ss, *Q = P + ss;
    const Cx<double> *bi = b + i * s, *xl = x + (i - h) * s;
    Cx<double> acc = {0, 0};
    for (int c = 0; c < s; c++) { cfma(acc, D[(i64)c * s], bi[c]); cfms(acc, P[(i64)c * s], xl[c]); }
    if (m + 1 < M) {
        const Cx<double> *xr = x + (i + h) * s;
        for (int c = 0; c < s; c++) cfms(acc, Q[(i64)c * s], xr[c]);
    }
    x[i * s + r] = acc;
}

// final block: x_0 = Dinv b_0
__global__ void k_bcr_final(int s, const Cx<double> *__restrict__ D, const Cx<double> *__restrict__ b,
                            Cx<double> *x)
{
    const int r = threadIdx.x;
    if (r >= s) return;
    Cx<double> acc = {0, 0};
    for (int c = 0; c < s; c++) cfma(acc, D[(i64)c * s + r], b[c]);
    x[r] = acc;
}

// border: t = b2 - dots; y2 = Sinv t -> y2out (ntau) and yout[n1 + a]
__global__ void k_border_small(int ntau, const Cx<double> *__restrict__ b2, const Cx<double> *__restrict__ dots,
                               const Cx<double> *__restrict__ Sinv, Cx<double> *y2out, Cx<double> *ytail)
{
    extern __shared__ char sm[];
    Cx<double> *t = reinterpret_cast<Cx<double> *>(sm);
    const int a = threadIdx.x;
    if (a < ntau) { t[a].re = b2[a].re - dots[a].re; t[a].im = b2[a].im - dots[a].im; }
    __syncthreads();
    if (a < ntau) {
        Cx<double> acc = {0, 0};
        for (int c = 0; c < ntau; c++) cfma(acc, Sinv[(i64)c * ntau + a], t[c]);
        y2out[a] = acc;
        ytail[a] = acc;
    }
}

// y[r] = xb[r] - sum_a Y[a n1 + r] y2[a], r < n1
__global__ __launch_bounds__(TPB) void k_border_update(i64 n1, int ntau, const Cx<double> *__restrict__ xb,
                                                       const Cx<double> *__restrict__ Y,
                                                       const Cx<double> *__restrict__ y2, Cx<double> *y)
{
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n1; i += (i64)gridDim.x * TPB) {
        Cx<double> acc = xb[i];
        for (int a = 0; a < ntau; a++) cfms(acc, Y[(i64)a * n1 + i], y2[a]);
        y[i] = acc;
    }
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
        hipLaunchKernelGGL((k_band_mv<float>), dim3(blocks_for(n1)), dim3(TPB), 0, s, (i64)n1, (i64)nb, bs, hw,
                           (const Cx<float> *)band, (const Cx<float> *)x, (Cx<float> *)y);
    else
        hipLaunchKernelGGL((k_band_mv<double>), dim3(blocks_for(n1)), dim3(TPB), 0, s, (i64)n1, (i64)nb, bs, hw,
                           (const Cx<double> *)band, (const Cx<double> *)x, (Cx<double> *)y);
    return chk();
}

int bbjd_o_diag_mv(hipStream_t s, int64_t n, const void *dg, const void *x, void *y, int single)
{
    if (single)
        hipLaunchKernelGGL((k_diag_mv<float>), dim3(gs_blocks(n)), dim3(TPB), 0, s, (i64)n, (const Cx<float> *)dg,
                           (const Cx<float> *)x, (Cx<float> *)y);
    else
        hipLaunchKernelGGL((k_diag_mv<double>), dim3(gs_blocks(n)), dim3(TPB), 0, s, (i64)n, (const Cx<double> *)dg,
                           (const Cx<double> *)x, (Cx<double> *)y);
    return chk();
}

int bbjd_o_bj_apply(hipStream_t s, int64_t n, int bs, const void *inv, const void *x, void *y, int single)
{
    if (single)
        hipLaunchKernelGGL((k_bj_apply<float>), dim3(blocks_for(n)), dim3(TPB), 0, s, (i64)n, bs,
                           (const Cx<float> *)inv, (const Cx<float> *)x, (Cx<float> *)y);
    else
        hipLaunchKernelGGL((k_bj_apply<double>), dim3(blocks_for(n)), dim3(TPB), 0, s, (i64)n, bs,
                           (const Cx<double> *)inv, (const Cx<double> *)x, (Cx<double> *)y);
    return chk();
}

int bbjd_o_bcr_fwd(hipStream_t st, int s, int64_t M, int64_t h, int64_t ne, const void *AG, void *b)
{
    if (ne <= 0) return 0;
    hipLaunchKernelGGL(k_bcr_fwd, dim3(blocks_for(ne * s)), dim3(TPB), 0, st, s, (i64)M, (i64)h, (i64)ne,
                       (const Cx<double> *)AG, (Cx<double> *)b);
    return chk();
}

int bbjd_o_bcr_bwd(hipStream_t st, int s, int64_t M, int64_t h, int64_t no, const void *DPQ, const void *b, void *x)
{
    if (no <= 0) return 0;
    hipLaunchKernelGGL(k_bcr_bwd, dim3(blocks_for(no * s)), dim3(TPB), 0, st, s, (i64)M, (i64)h, (i64)no,
                       (const Cx<double> *)DPQ, (const Cx<double> *)b, (Cx<double> *)x);
    return chk();
}

int bbjd_o_bcr_final(hipStream_t st, int s, const void *D, const void *b, void *x)
{
    hipLaunchKernelGGL(k_bcr_final, dim3(1), dim3(s <= 64 ? 64 : 256), 0, st, s, (const Cx<double> *)D,
                       (const Cx<double> *)b, (Cx<double> *)x);
    return chk();
}

int bbjd_o_border_small(hipStream_t st, int ntau, const void *b2, const void *dots, const void *Sinv, void *y2,
                        void *ytail)
{
    hipLaunchKernelGGL(k_border_small, dim3(1), dim3(ntau <= 64 ? 64 : 256), (size_t)ntau * 16, st, ntau,
                       (const Cx<double> *)b2, (const Cx<double> *)dots, (const Cx<double> *)Sinv,
                       (Cx<double> *)y2, (Cx<double> *)ytail);
    return chk();
}

int bbjd_o_border_update(hipStream_t st, int64_t n1, int ntau, const void *xb, const void *Y, const void *y2,
                         void *y)
{
    hipLaunchKernelGGL(k_border_update, dim3(gs_blocks(n1)), dim3(TPB), 0, st, (i64)n1, ntau,
                       (const Cx<double> *)xb, (const Cx<double> *)Y, (const Cx<double> *)y2, (Cx<double> *)y);
    return chk();
}

}  // extern "C"
