// kernels.hip -- streaming kernels of the JDQZ outer loop for gfx950.
//
// Every n-sized step of one outer iteration (reference src/jd_core.c:
// CGS2 207-241, project_out 269-309, Ritz vector 870-877, residual 887-892,
// Sleijpen-van der Vorst step 952-971, pencil update 1085-1104, thick
// restart 975-1030) is HBM-bound tall-skinny work on column-major tiles.
//
// Design for CDNA (64-wide wavefronts):
//  * projections Q^H x: a block owns 1024 rows x C columns; each thread
//    keeps 4 rows of x in registers and accumulates 2*NV*C = 32 real
//    partial sums, reduced across the wavefront by a transposing register
//    butterfly (31 lane exchanges for 32 values) instead of 32 separate
//    6-step reductions, then across the 4 waves through 512 B of LDS;
//  * second stage sums tile partials in a fixed order -> results do not
//    depend on scheduling or on what else runs on the GPU;
//  * updates x += alpha Q h stage h in LDS (broadcast reads) and fuse the
//    squared norm of the result;
//  * the DGKS decision (second pass iff ||v_after|| < 0.7071 ||v_before||)
//    is taken on the device: pass-2 kernels read a flag and exit early, so
//    an orthogonalisation needs no host round trip.
//  * the same templates give the complex64 kernels (FP32 accumulation) of
//    the mixed-precision inner solve.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstring>
#include <mutex>
#include <vector>
#include "engine.h"
#include "launch.h"

typedef long long i64;

// ---------------------------------------------------------------------------
// Launch timing of the projection kernel (k_proj, one vector, first CGS
// pass): when enabled, every launch carries a start and a stop HIP event in
// its dispatch (hipExtLaunchKernelGGL), so their interval is the kernel's
// execution, and its algorithmic bytes 16 n (k + 1) are recorded;
// bbjd_ktime_harvest() turns the events into durations.
namespace {
struct KRec { hipEvent_t a, b; double bytes; };
std::mutex g_kt_mu;
std::vector<KRec> g_kt;
std::vector<hipEvent_t> g_kt_pool;
bool g_kt_on = false;
constexpr size_t kKtMaxRecords = 1 << 16;   // launches timed between harvests
hipEvent_t kt_event()
{
    if (!g_kt_pool.empty()) { hipEvent_t e = g_kt_pool.back(); g_kt_pool.pop_back(); return e; }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}
}  // namespace

template <typename R> struct alignas(2 * sizeof(R)) Cx { R re, im; };

template <typename R> __device__ __forceinline__ Cx<R> cx(R a, R b) { Cx<R> c; c.re = a; c.im = b; return c; }
template <typename R> __device__ __forceinline__ Cx<R> operator+(Cx<R> a, Cx<R> b) { return cx(a.re + b.re,
                                                                 a.im + b.im); }
template <typename R> __device__ __forceinline__ Cx<R> operator-(Cx<R> a, Cx<R> b) { return cx(a.re - b.re,
                                                                 a.im - b.im); }
template <typename R> __device__ __forceinline__ Cx<R> operator*(Cx<R> a, Cx<R> b)
{ return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
// Accumulations with an explicit fused multiply-add of one product and the
// other product rounded, then added to the accumulator: the fusion does not
// depend on the compiler's choices, and the accumulator carries one
// dependent add per term (a chain of two fused multiply-adds into it measured
// 5% slower in the projection at k = 128).
__device__ __forceinline__ double ffma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float ffma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// a += conj(q) * x
template <typename R> __device__ __forceinline__ void cfma_conj(R &are, R &aim, Cx<R> q, Cx<R> x)
{
    are += ffma(q.re, x.re, q.im * x.im);
    aim += ffma(q.re, x.im, -(q.im * x.re));
}
// a += q * h
template <typename R> __device__ __forceinline__ void cfma(Cx<R> &a, Cx<R> q, Cx<R> h)
{
    a.re += ffma(q.re, h.re, -(q.im * h.im));
    a.im += ffma(q.re, h.im, q.im * h.re);
}

static constexpr int TPB = BBJD_TPB;
typedef double v4d __attribute__((ext_vector_type(4)));   // f64 MFMA accumulator

// Reduce 2^LS values per lane over the 64-lane wavefront.  Afterwards lane l
// holds in v[0] the wave sum of value index (l >> (6 - LS)).
template <typename R, int LS>
__device__ __forceinline__ void wave_transpose_reduce(R (&v)[1 << LS])
{
    constexpr int S = 1 << LS;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int st = 0; st < LS; st++) {
        const int h = S >> (st + 1);
        const int m = 32 >> st;
        const bool up = (lane & m) != 0;
#pragma unroll
        for (int i = 0; i < h; i++) {
            // select between values, not addresses: a select of two array
            // elements can become a dynamically indexed (scratch) array
            const R lo = v[i], hi = v[i + h];
            const R send = up ? lo : hi;
            const R keep = up ? hi : lo;
            v[i] = keep + __shfl_xor(send, m, 64);
        }
    }
#pragma unroll
    for (int m = (32 >> LS); m >= 1; m >>= 1) v[0] += __shfl_xor(v[0], m, 64);
}

template <typename R> __device__ __forceinline__ R wave_sum(R v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Block-wide sum of one value per thread (TPB = 256 -> 4 waves); result valid
// in thread 0.  Fixed order.
template <typename R> __device__ __forceinline__ R block_sum(R v, R *lds4)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    R s = 0;
    if (threadIdx.x == 0) s = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    return s;
}

// ---------------------------------------------------------------------------
// Last-block-done finishing.  Every reduction kernel writes per-block
// partials; the last block of a group to arrive (counter in bbjd_red.cnt)
// sums them in a fixed order -- tile 0, 1, 2, ... -- so the result does not
// depend on scheduling, and resets the counter for the next launch on the
// stream.  Only thread 0 of a block fences: an agent-scope release (L2
// write-back, after the workgroup barrier has ordered the block's stores)
// before the counter increment, and, in the last block only, an agent-scope
// acquire (L2 invalidate) before the partials of blocks on other XCDs are
// read.  A fence in every thread would write back and invalidate the L2 once
// per thread group and throttle the streaming kernels.
__device__ __forceinline__ bool last_block_of(int *cnt, int group_size, int *lds_flag)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const int last = atomicAdd(cnt, 1) == group_size - 1;
        if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *lds_flag = last;
    }
    __syncthreads();
    return *lds_flag != 0;
}

// What the last block of a norm-producing kernel does with the norm.
//  mode 0: scal[slot] = ||x||
//  mode 1: after CGS pass 1: also scal[SL_NA] = ||x||, flag[FL_PASS2] =
//          ||x|| < 0.7071 scal[SL_NB]; hsum[:k] = h when hsum
//  mode 2: after CGS pass 2 (runs only when flagged): scal[slot] = ||x||,
//          hsum[:k] = h + h2 when hsum
//  mode 3: after the second pass of the fused CGS2 with unit output:
//          scal[slot] = ||x||, and x /= ||x|| by this block when the pass
//          left the scaling to it (flag[FL_RESCALE]; rare)
struct NormFin {
    double *scal;
    int *flag;
    int mode, slot, k;
    void *hsum;
    const void *h, *h2;
    void *x;             // mode 3: the vector to rescale (n entries)
    i64 n;
};

template <typename R>
__device__ __forceinline__ void finish_norm(const R *pnorm, int nblk, const NormFin &f, R *lds4)
{
    __shared__ double lsc;
    R s = 0;
    for (int t = threadIdx.x; t < nblk; t += TPB) s += pnorm[t];
    s = block_sum(s, lds4);
    if (threadIdx.x == 0) {
        const double nrm = sqrt((double)s);
        f.scal[f.slot] = nrm;
        if (f.mode == 1) {
            f.scal[SL_NA] = nrm;
            const int p2 = nrm < 0.7071 * f.scal[SL_NB];
            f.flag[FL_PASS2] = p2;
            if (p2) f.flag[FL_DGKS_COUNT] += 1;
        }
        if (f.mode == 3) lsc = (f.flag[FL_RESCALE] && nrm > 1e-300) ? 1.0 / nrm : 0.0;
    }
    if (f.mode == 3) {
        __syncthreads();
        const R sc = (R)lsc;
        if (lsc != 0.0) {
            Cx<R> *x = reinterpret_cast<Cx<R> *>(f.x);
            for (i64 i = threadIdx.x; i < f.n; i += TPB) {
                Cx<R> v = x[i];
                v.re *= sc;
                v.im *= sc;
                x[i] = v;
            }
        }
    }
    if (f.hsum) {
        Cx<R> *hs = reinterpret_cast<Cx<R> *>(f.hsum);
        const Cx<R> *h = reinterpret_cast<const Cx<R> *>(f.h), *h2 = reinterpret_cast<const Cx<R> *>(f.h2);
        for (int j = threadIdx.x; j < f.k; j += TPB) hs[j] = f.mode == 2 ? h[j] + h2[j] : h[j];
    }
}

// Projection Q^H x (and Q^H x1; with three vectors Q^H [x, x1, x2] for
// x = ca x1 + cb x2, formed in registers): a block owns a chunk of rows x C columns
// and writes its partial sums part[(tile*NV + vi)*ldp + col]; the last row
// tile of each column tile to finish reduces that column tile's partials into
// out0/out1.  With want_norm the column-tile-0 blocks also produce ||x0||^2
// partials and the last of them writes scal[norm_slot] = ||x0||.
// grid = (column tiles, row chunks).
//
// Finish of a projection for column tile c0: the partials of ntiles row
// tiles in a fixed order (16 lanes per (vector, column) pair), plus the norm
// when norm_here.  Run by the last block of the column tile (fused) or by
// k_proj_finish.
template <typename R, int NV, int C>
__device__ __forceinline__ void proj_finish(int c0, int ntiles, int k, const Cx<R> *part, int ldp, const R *pnorm,
                                            bool norm_here, Cx<R> *out0, Cx<R> *out1, Cx<R> *out2, double *scal,
                                            int norm_slot, R *lds4)
{
    constexpr int PAIRS = NV * C;                 // (vector, column) pairs
    constexpr int LP = PAIRS <= 16 ? 16 : 8;      // lanes per pair
    static_assert(PAIRS * LP <= TPB, "finish lanes");
    const int t = threadIdx.x;
    const int g = t / LP, l = t % LP;
    const int vi = g / C, c = g % C, col = c0 + c;
    const bool mine = g < PAIRS && col < k;
    R sr = 0, si = 0;
    if (mine)
        for (int tt = l; tt < ntiles; tt += LP) {
            const Cx<R> p = part[((i64)tt * NV + vi) * ldp + col];
            sr += p.re;
            si += p.im;
        }
#pragma unroll
    for (int m = LP / 2; m >= 1; m >>= 1) {
        sr += __shfl_xor(sr, m, 64);
        si += __shfl_xor(si, m, 64);
    }
    if (l == 0 && mine) (vi == 0 ? out0 : vi == 1 ? out1 : out2)[col] = cx(sr, si);
    if (norm_here) {
        R s = 0;
        for (int tt = t; tt < ntiles; tt += TPB) s += pnorm[tt];
        s = block_sum(s, lds4);
        if (t == 0) scal[norm_slot] = sqrt((double)s);
    }
}

// COND: the launch is the conditional second CGS pass (a separate
// instantiation, so profiles tell the passes that ran from those that
// returned at once).
template <typename R, int NV, int C, bool COND>
__device__ __forceinline__ void k_proj_b(const VB vb, i64 n, i64 chunk, int k, const Cx<R> *__restrict__ Q, i64 ldq,
                                              const Cx<R> *__restrict__ x0, const Cx<R> *__restrict__ x1,
                                              const Cx<R> *__restrict__ x2, Cx<R> ca, Cx<R> cb, Cx<R> *part,
                                              int ldp, R *pnorm, int want_norm, const int *__restrict__ cond,
                                              int *cnt, Cx<R> *out0, Cx<R> *out1, Cx<R> *out2, double *scal,
                                              int norm_slot, int fused)
{
    if (COND && cond[0] == 0) return;
    constexpr int NVAL = 2 * NV * C;
    static_assert(NVAL == 32 || NVAL == 16 || NVAL == 48, "butterfly width");
    constexpr int LS = NVAL == 16 ? 4 : 5;        // 48 values: a 32 and a 16 butterfly
    __shared__ R lds[4][NVAL];
    __shared__ R lds4[4];
    __shared__ int lflag;
    const int t = threadIdx.x;
    const int tile = vb.by;
    const int c0 = vb.bx * C;
    const i64 r1 = (i64)(tile + 1) * chunk < n ? (i64)(tile + 1) * chunk : n;
    const bool do_norm = want_norm && vb.bx == 0;
    const Cx<R> *q0 = Q + (i64)c0 * ldq;

    // A block streams its chunk of rows; a thread accumulates C columns
    // (NV vectors) over rows t, t + TPB, ... of the chunk.
    R acc[NVAL];
#pragma unroll
    for (int i = 0; i < NVAL; i++) acc[i] = 0;
    R nv = 0;
    // A partial last column tile reads its last column again in place of
    // the missing ones (cache hits; their sums are discarded), so every tile
    // runs the same unrolled loop.
    const int ncol = k - c0 < C ? k - c0 : C;
    const Cx<R> *qp[C];
#pragma unroll
    for (int c = 0; c < C; c++) qp[c] = q0 + (i64)(c < ncol ? c : ncol - 1) * ldq;
#pragma unroll 2
    for (i64 r = (i64)tile * chunk + t; r < r1; r += TPB) {
        Cx<R> xv[NV];
        if (NV == 3) {
            // three vectors: x1, x2 and their combination ca x1 + cb x2
            xv[1] = x1[r];
            xv[NV - 1] = x2[r];
            xv[0] = ca * xv[1] + cb * xv[NV - 1];
        } else {
            xv[0] = x0[r];
            if (NV == 2) xv[NV - 1] = x1[r];
        }
        Cx<R> qv[C];
#pragma unroll
        for (int c = 0; c < C; c++) qv[c] = qp[c][r];
        if (do_norm) nv += xv[0].re * xv[0].re + xv[0].im * xv[0].im;
#pragma unroll
        for (int c = 0; c < C; c++)
#pragma unroll
            for (int vi = 0; vi < NV; vi++)
                cfma_conj(acc[2 * (vi * C + c)], acc[2 * (vi * C + c) + 1], qv[c], xv[vi]);
    }
    const int lane = t & 63, w = t >> 6;
    if constexpr (NVAL == 48) {
        // values 0..31 and 32..47 reduced separately; lane l then holds
        // value l >> 1 of the first group and 32 + (l >> 2) of the second
        R(&a32)[32] = *reinterpret_cast<R(*)[32]>(&acc[0]);
        R(&a16)[16] = *reinterpret_cast<R(*)[16]>(&acc[NVAL - 16]);
        wave_transpose_reduce<R, 5>(a32);
        wave_transpose_reduce<R, 4>(a16);
        if ((lane & 1) == 0) lds[w][lane >> 1] = a32[0];
        if ((lane & 3) == 0) lds[w][32 + (lane >> 2)] = a16[0];
    } else {
        R(&av)[1 << LS] = *reinterpret_cast<R(*)[1 << LS]>(&acc[0]);
        wave_transpose_reduce<R, LS>(av);
        // after the butterfly, lane l holds value index l >> (6 - LS)
        if ((lane & ((64 >> LS) - 1)) == 0) lds[w][lane >> (6 - LS)] = av[0];
    }
    __syncthreads();
    if (t < NVAL) {
        const R s = (lds[0][t] + lds[1][t]) + (lds[2][t] + lds[3][t]);
        const int vi = t / (2 * C), c = (t / 2) % C, part_im = t & 1;
        const int col = c0 + c;
        if (col < k) {
            R *dst = reinterpret_cast<R *>(part + ((i64)tile * NV + vi) * ldp + col);
            dst[part_im] = s;
        }
    }
    if (do_norm) {
        const R s = block_sum(nv, lds4);
        if (t == 0) pnorm[tile] = s;
    }
    if (!fused) return;
    if (!last_block_of(cnt + vb.bx, vb.gy, &lflag)) return;
    proj_finish<R, NV, C>(c0, vb.gy, k, part, ldp, pnorm, want_norm && vb.bx == 0, out0, out1, out2,
                          scal, norm_slot, lds4);
    if (t == 0) cnt[vb.bx] = 0;
}
template <typename R, int NV, int C, bool COND>
__global__ __launch_bounds__(TPB) void k_proj(i64 n, i64 chunk, int k, const Cx<R> *__restrict__ Q, i64 ldq,
                                              const Cx<R> *__restrict__ x0, const Cx<R> *__restrict__ x1,
                                              const Cx<R> *__restrict__ x2, Cx<R> ca, Cx<R> cb, Cx<R> *part,
                                              int ldp, R *pnorm, int want_norm, const int *__restrict__ cond,
                                              int *cnt, Cx<R> *out0, Cx<R> *out1, Cx<R> *out2, double *scal,
                                              int norm_slot, int fused)
{
    k_proj_b<R, NV, C, COND>(VB_HW, n, chunk, k, Q, ldq, x0, x1, x2, ca, cb, part, ldp, pnorm, want_norm, cond, cnt,
                             out0, out1, out2, scal, norm_slot, fused);
}

template <typename R, int NV, int C>
__device__ __forceinline__ void k_proj_finish_b(const VB vb, int ntiles, int k, const Cx<R> *part, int ldp,
                                                const R *pnorm,
                                                     int want_norm, const int *__restrict__ cond, Cx<R> *out0,
                                                     Cx<R> *out1, Cx<R> *out2, double *scal, int norm_slot)
{
    if (cond && cond[0] == 0) return;
    __shared__ R lds4[4];
    proj_finish<R, NV, C>(vb.bx * C, ntiles, k, part, ldp, pnorm, want_norm && vb.bx == 0, out0, out1,
                          out2, scal, norm_slot, lds4);
}
template <typename R, int NV, int C>
__global__ __launch_bounds__(TPB) void k_proj_finish(int ntiles, int k, const Cx<R> *part, int ldp, const R *pnorm,
                                                     int want_norm, const int *__restrict__ cond, Cx<R> *out0,
                                                     Cx<R> *out1, Cx<R> *out2, double *scal, int norm_slot)
{
    k_proj_finish_b<R, NV, C>(VB_HW, ntiles, k, part, ldp, pnorm, want_norm, cond, out0, out1, out2, scal, norm_slot);
}

// x = beta*x + alpha * Q[:, :k] h; with fin.mode >= 0 the squared norm of
// the result is reduced (last block) and finished per NormFin.
// Dynamic LDS: k complex.
template <typename R, int E>
__device__ __forceinline__ void k_update_b(const VB vb, i64 n, int k, const Cx<R> *__restrict__ Q, i64 ldq,
                                                const Cx<R> *__restrict__ h, R alpha, int beta_one,
                                                Cx<R> *__restrict__ x, R *pnorm, const int *__restrict__ cond,
                                                int *cnt, NormFin fin, int fused)
{
    if (cond && cond[0] == 0) return;
    extern __shared__ char smem_raw[];
    Cx<R> *sh = reinterpret_cast<Cx<R> *>(smem_raw);
    __shared__ R lds4[4];
    __shared__ int lflag;
    const int t = threadIdx.x;
    for (int j = t; j < k; j += TPB) { const Cx<R> hj = h[j]; sh[j] = cx(alpha * hj.re, alpha * hj.im); }
    __syncthreads();
    const i64 rb = (i64)vb.bx * (TPB * E);
    Cx<R> a[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const i64 r = rb + t + (i64)e * TPB;
        a[e] = (beta_one && r < n) ? x[r] : cx<R>(0, 0);
    }
    if (rb + TPB * E <= n) {
        constexpr int UNR = E == 4 ? 4 : 8;     // loads in flight per thread
#pragma unroll UNR
        for (int j = 0; j < k; j++) {
            const Cx<R> hj = sh[j];
            const Cx<R> *q = Q + (i64)j * ldq + rb + t;
#pragma unroll
            for (int e = 0; e < E; e++) cfma(a[e], q[(i64)e * TPB], hj);
        }
    } else {
        for (int j = 0; j < k; j++) {
            const Cx<R> hj = sh[j];
            const Cx<R> *q = Q + (i64)j * ldq;
#pragma unroll
            for (int e = 0; e < E; e++) {
                const i64 r = rb + t + (i64)e * TPB;
                if (r < n) cfma(a[e], q[r], hj);
            }
        }
    }
    R s = 0;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const i64 r = rb + t + (i64)e * TPB;
        if (r < n) { x[r] = a[e]; s += a[e].re * a[e].re + a[e].im * a[e].im; }
    }
    if (fin.mode < 0) return;
    s = block_sum(s, lds4);
    if (t == 0) pnorm[vb.bx] = s;
    if (!fused || !last_block_of(cnt, vb.gx, &lflag)) return;
    finish_norm<R>(pnorm, vb.gx, fin, lds4);
    if (t == 0) *cnt = 0;
}
template <typename R, int E>
__global__ __launch_bounds__(TPB) void k_update(i64 n, int k, const Cx<R> *__restrict__ Q, i64 ldq,
                                                const Cx<R> *__restrict__ h, R alpha, int beta_one,
                                                Cx<R> *__restrict__ x, R *pnorm, const int *__restrict__ cond,
                                                int *cnt, NormFin fin, int fused)
{
    k_update_b<R, E>(VB_HW, n, k, Q, ldq, h, alpha, beta_one, x, pnorm, cond, cnt, fin, fused);
}

// x -= Y[:, :k] (M c) with a k x k matrix M (column-major) applied to the
// coefficients first: every block forms M c in LDS (k^2 products, M read from
// cache), then streams its rows of Y.  The inner solve's oblique projector
// (I - Y H^-1 Q^H) uses it with M = H^-1.  Dynamic LDS: 2 k complex.
template <typename R, int E>
__global__ __launch_bounds__(TPB) void k_update_pre(i64 n, int k, const Cx<R> *__restrict__ Y, i64 ldy,
                                                    const Cx<R> *__restrict__ M, const Cx<R> *__restrict__ c,
                                                    Cx<R> *__restrict__ x)
{
    extern __shared__ char smem_raw[];
    Cx<R> *sc = reinterpret_cast<Cx<R> *>(smem_raw);
    Cx<R> *sh = sc + k;
    const int t = threadIdx.x;
    for (int j = t; j < k; j += TPB) sc[j] = c[j];
    __syncthreads();
    for (int j = t; j < k; j += TPB) {
        Cx<R> a = cx<R>(0, 0);
        for (int i = 0; i < k; i++) cfma(a, M[(i64)i * k + j], sc[i]);
        sh[j] = cx(-a.re, -a.im);
    }
    __syncthreads();
    const i64 rb = (i64)blockIdx.x * (TPB * E);
    Cx<R> a[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const i64 r = rb + t + (i64)e * TPB;
        a[e] = r < n ? x[r] : cx<R>(0, 0);
    }
    if (rb + TPB * E <= n) {
#pragma unroll 4
        for (int j = 0; j < k; j++) {
            const Cx<R> hj = sh[j];
            const Cx<R> *q = Y + (i64)j * ldy + rb + t;
#pragma unroll
            for (int e = 0; e < E; e++) cfma(a[e], q[(i64)e * TPB], hj);
        }
    } else {
        for (int j = 0; j < k; j++) {
            const Cx<R> hj = sh[j];
            const Cx<R> *q = Y + (i64)j * ldy;
#pragma unroll
            for (int e = 0; e < E; e++) {
                const i64 r = rb + t + (i64)e * TPB;
                if (r < n) cfma(a[e], q[r], hj);
            }
        }
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
        const i64 r = rb + t + (i64)e * TPB;
        if (r < n) x[r] = a[e];
    }
}

// The same update with the columns split over the block's four waves: a
// block owns 64 * RW rows, wave w accumulates columns w, w + 4, ... of its
// rows, and the four partial sums are added in a fixed order through LDS.
// Four times the blocks of k_update for the same rows, so more loads are in
// flight on shapes whose rows alone do not fill the GPU.
template <typename R, int RW, bool COND>
__device__ __forceinline__ void k_update_split_b(const VB vb, i64 n, int k, const Cx<R> *__restrict__ Q, i64 ldq,
                                                      const Cx<R> *__restrict__ h, R alpha, int beta_one,
                                                      Cx<R> *__restrict__ x, R *pnorm,
                                                      const int *__restrict__ cond, int *cnt, NormFin fin,
                                                      int fused)
{
    if (COND && cond[0] == 0) return;
    extern __shared__ char smem_raw[];
    Cx<R> *sh = reinterpret_cast<Cx<R> *>(smem_raw);
    __shared__ Cx<R> red[4][64 * RW];
    __shared__ R lds4[4];
    __shared__ int lflag;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int j = t; j < k; j += TPB) { const Cx<R> hj = h[j]; sh[j] = cx(alpha * hj.re, alpha * hj.im); }
    __syncthreads();
    const i64 rb = (i64)vb.bx * (64 * RW);
    Cx<R> a[RW];
#pragma unroll
    for (int e = 0; e < RW; e++) a[e] = cx<R>(0, 0);
    if (rb + 64 * RW <= n) {
#pragma unroll 8
        for (int j = w; j < k; j += 4) {
            const Cx<R> hj = sh[j];
            const Cx<R> *q = Q + (i64)j * ldq + rb + lane;
#pragma unroll
            for (int e = 0; e < RW; e++) cfma(a[e], q[e * 64], hj);
        }
    } else {
        for (int j = w; j < k; j += 4) {
            const Cx<R> hj = sh[j];
            const Cx<R> *q = Q + (i64)j * ldq;
#pragma unroll
            for (int e = 0; e < RW; e++) {
                const i64 r = rb + lane + e * 64;
                if (r < n) cfma(a[e], q[r], hj);
            }
        }
    }
#pragma unroll
    for (int e = 0; e < RW; e++) red[w][lane + e * 64] = a[e];
    __syncthreads();
    R s = 0;
    if (t < 64 * RW) {
        const i64 r = rb + t;
        if (r < n) {
            Cx<R> v = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
            if (beta_one) v = x[r] + v;
            x[r] = v;
            s = v.re * v.re + v.im * v.im;
        }
    }
    if (fin.mode < 0) return;
    s = block_sum(s, lds4);
    if (t == 0) pnorm[vb.bx] = s;
    if (!fused || !last_block_of(cnt, vb.gx, &lflag)) return;
    finish_norm<R>(pnorm, vb.gx, fin, lds4);
    if (t == 0) *cnt = 0;
}
template <typename R, int RW, bool COND>
__global__ __launch_bounds__(TPB) void k_update_split(i64 n, int k, const Cx<R> *__restrict__ Q, i64 ldq,
                                                      const Cx<R> *__restrict__ h, R alpha, int beta_one,
                                                      Cx<R> *__restrict__ x, R *pnorm,
                                                      const int *__restrict__ cond, int *cnt, NormFin fin,
                                                      int fused)
{
    k_update_split_b<R, RW, COND>(VB_HW, n, k, Q, ldq, h, alpha, beta_one, x, pnorm, cond, cnt, fin, fused);
}

// ---------------------------------------------------------------------------
// Fused CGS2.  When the DGKS test asks for the second pass, CGS2 reads Q four
// times (projection and update of each pass).  The fused form reads it three
// times:
//   k_proj        h1 = Q^H v0 and ||v0||                            read 1
//   k_cgs_mid     v1 = v0 - Q h1, partials of Q^H v1 and ||v1||^2   read 2
//   (finish)      h2 = Q^H v1, ||v1||, the DGKS decision
//   k_cgs_final   v2 = v1 - Q h2 and ||v2|| when decided            read 3
// A block of k_cgs_mid owns 64-row tiles, one row per lane; wave w holds the
// tile's entries of columns w, w + 4, ... in registers, so the update and the
// projection of a row use the same loaded values.  The projection is computed
// for every call, also when the test then passes (it costs no extra reads).
// Unit-norm output folds the normalisation into the last launch: without the
// second pass v1 / ||v1||; with it v2 / sqrt(||v1||^2 - ||h2||^2), which is
// v2 / ||v2|| to rounding while ||h2||^2 <= 3/4 ||v1||^2 (Q orthonormal);
// otherwise v2 stays unscaled, FL_RESCALE is set, and the block that finishes
// ||v2|| scales it (rare: v0 in span(Q) to a few ulp).
struct MidFin {
    double *scal;
    int *flag;
    int out_slot;
    int decide;          // 1: CGS2 (norm, DGKS decision, flags); 0: projection only
    const void *h1;      // first-pass coefficients (for hsum)
    void *h2, *hsum;
};

// h2 from the k_cgs_mid partials (columns col0, col0 + col_step, ... one wave
// each, fixed order), and with decide the norm, the decision and the flags
// (written by the writer block only; every finishing block computes them).
template <typename R>
__device__ __forceinline__ void cgs_mid_finish(int col0, int col_step, bool writer, int nblk, int k,
                                               const Cx<R> *part, const R *pnorm, const MidFin &f, R *lds4,
                                               int *lds_p2)
{
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int p2 = 0;
    if (f.decide) {
        R s = 0;
        for (int b = t; b < nblk; b += TPB) s += pnorm[b];
        s = block_sum(s, lds4);
        if (t == 0) {
            const double nrm = sqrt((double)s);
            const int pass2 = nrm < 0.7071 * f.scal[SL_NB];
            *lds_p2 = pass2;
            if (writer) {
                f.scal[SL_NA] = nrm;
                f.scal[f.out_slot] = nrm;
                f.flag[FL_PASS2] = pass2;
                f.flag[FL_RESCALE] = 0;
                if (pass2) f.flag[FL_DGKS_COUNT] += 1;
            }
        }
        __syncthreads();
        p2 = *lds_p2;
    }
    Cx<R> *h2 = reinterpret_cast<Cx<R> *>(f.h2);
    Cx<R> *hs = reinterpret_cast<Cx<R> *>(f.hsum);
    const Cx<R> *h1 = reinterpret_cast<const Cx<R> *>(f.h1);
    for (int j = col0 + w; j < k; j += col_step) {
        R sr = 0, si = 0;
        for (int b = lane; b < nblk; b += 64) {
            const Cx<R> p = part[(i64)b * k + j];
            sr += p.re;
            si += p.im;
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        if (lane == 0) {
            h2[j] = cx(sr, si);
            if (hs) hs[j] = p2 ? h1[j] + cx(sr, si) : h1[j];
        }
    }
}

// v1 = v0 - Q h1 (in place; with xa, v0 = ca xa + cb xb and v1 goes to x)
// and the block's partials of Q^H v1 (part[block k + j]) and, with
// fin.decide, of ||v1||^2.  KW >= ceil(k / 4) columns per wave; a block walks
// `tiles` consecutive 64-row tiles.
template <typename R, int KW>
__device__ __forceinline__ void k_cgs_mid_b(const VB vb, i64 n, int k, const Cx<R> *__restrict__ Q, i64 ldq,
                                                 const Cx<R> *__restrict__ h, Cx<R> *x, const Cx<R> *__restrict__ xa,
                                                 const Cx<R> *__restrict__ xb, Cx<R> ca, Cx<R> cb, int tiles,
                                                 Cx<R> *part, R *pnorm, int *cnt, MidFin fin, int fused)
{
    constexpr int NVAL = 2 * KW;                  // projection values per lane
    constexpr int G = NVAL < 16 ? NVAL : 16;      // values per butterfly
    constexpr int NG = NVAL / G;
    constexpr int LS = G == 16 ? 4 : (G == 8 ? 3 : 2);
    static_assert((1 << LS) == G, "butterfly width");
    __shared__ Cx<R> sh[4 * KW];
    __shared__ Cx<R> red[4][64];
    __shared__ Cx<R> xin[64];
    __shared__ R lds4[4];
    __shared__ int lflag, lp2;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int j = t; j < 4 * KW; j += TPB) sh[j] = j < k ? h[j] : cx<R>(0, 0);
    __syncthreads();
    R acc[NG];
#pragma unroll
    for (int g = 0; g < NG; g++) acc[g] = 0;
    R nrm = 0;
    const i64 tb0 = (i64)vb.bx * tiles;
    for (int tt = 0; tt < tiles; tt++) {
        const i64 rb = (tb0 + tt) * 64;
        if (rb >= n) break;                       // uniform over the block
        const i64 r = rb + lane;
        const bool in = r < n;
        Cx<R> q[KW];
#pragma unroll
        for (int i = 0; i < KW; i++) {
            const int j = w + 4 * i;
            q[i] = (in && j < k) ? Q[(i64)j * ldq + r] : cx<R>(0, 0);
        }
        if (w == 0) xin[lane] = !in ? cx<R>(0, 0) : xa ? ca * xa[r] + cb * xb[r] : x[r];
        Cx<R> a = cx<R>(0, 0);
#pragma unroll
        for (int i = 0; i < KW; i++) cfma(a, q[i], sh[w + 4 * i]);
        red[w][lane] = a;
        __syncthreads();
        const Cx<R> v = xin[lane] - ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]));
        if (w == 0 && in) {
            x[r] = v;
            nrm += v.re * v.re + v.im * v.im;
        }
#pragma unroll
        for (int g = 0; g < NG; g++) {
            R vals[G];
#pragma unroll
            for (int c = 0; c < G / 2; c++) {
                const Cx<R> qc = q[g * (G / 2) + c];
                vals[2 * c] = qc.re * v.re + qc.im * v.im;          // conj(q) v
                vals[2 * c + 1] = qc.re * v.im - qc.im * v.re;
            }
            wave_transpose_reduce<R, LS>(vals);
            acc[g] += vals[0];
        }
        __syncthreads();                          // red, xin are rewritten next tile
    }
    // lane l holds value index l >> (6 - LS) of each butterfly group
    constexpr int SH = 6 - LS;
    if ((lane & ((1 << SH) - 1)) == 0) {
        const int vi = lane >> SH;
#pragma unroll
        for (int g = 0; g < NG; g++) {
            const int j = w + 4 * (g * (G / 2) + (vi >> 1));
            if (j < k) reinterpret_cast<R *>(part + (i64)vb.bx * k + j)[vi & 1] = acc[g];
        }
    }
    if (fin.decide && w == 0) {
        const R s = wave_sum(nrm);
        if (lane == 0) pnorm[vb.bx] = s;
    }
    if (!fused || !last_block_of(cnt, vb.gx, &lflag)) return;
    cgs_mid_finish<R>(0, 4, true, vb.gx, k, part, pnorm, fin, lds4, &lp2);
    if (t == 0) *cnt = 0;
}
template <typename R, int KW>
__global__ __launch_bounds__(TPB) void k_cgs_mid(i64 n, int k, const Cx<R> *__restrict__ Q, i64 ldq,
                                                 const Cx<R> *__restrict__ h, Cx<R> *x, const Cx<R> *__restrict__ xa,
                                                 const Cx<R> *__restrict__ xb, Cx<R> ca, Cx<R> cb, int tiles,
                                                 Cx<R> *part, R *pnorm, int *cnt, MidFin fin, int fused)
{
    k_cgs_mid_b<R, KW>(VB_HW, n, k, Q, ldq, h, x, xa, xb, ca, cb, tiles, part, pnorm, cnt, fin, fused);
}

template <typename R>
__device__ __forceinline__ void k_cgs_mid_fin_b(const VB vb, int nblk, int k, const Cx<R> *part, const R *pnorm,
                                                     MidFin fin)
{
    __shared__ R lds4[4];
    __shared__ int lp2;
    cgs_mid_finish<R>(vb.bx * 4, vb.gx * 4, vb.bx == 0, nblk, k, part, pnorm, fin, lds4, &lp2);
}
template <typename R>
__global__ __launch_bounds__(TPB) void k_cgs_mid_fin(int nblk, int k, const Cx<R> *part, const R *pnorm,
                                                     MidFin fin)
{
    k_cgs_mid_fin_b<R>(VB_HW, nblk, k, part, pnorm, fin);
}

// Last launch of the fused CGS2 (k <= 128).  With the second pass (decided,
// or forced for project_out): x = (x - Q h2) * sc with ||x - Q h2|| reduced
// per fin; sc = 1 / sqrt(||v1||^2 - ||h2||^2) for unit output when that is
// at least trust ||v1||^2 (trust 1/4: the estimate is then good to a few
// ulp), else 1 and FL_RESCALE.  Without it: x *= 1 / ||v1|| for unit
// output (when ||v1|| > 1e-300), else nothing.  Columns split over the four
// waves as in k_update_split.
template <typename R, int RW>
__device__ __forceinline__ void k_cgs_final_b(const VB vb, i64 n, int k, const Cx<R> *__restrict__ Q, i64 ldq,
                                                   const Cx<R> *__restrict__ h2, Cx<R> *__restrict__ x, int scale,
                                                   int force, double trust, const double *scal, int *flag,
                                                   R *pnorm, int *cnt, NormFin fin, int fused)
{
    __shared__ Cx<R> sh[128];
    __shared__ Cx<R> red[4][64 * RW];
    __shared__ R lds4[4];
    __shared__ int lflag;
    __shared__ double lsc;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const i64 rb = (i64)vb.bx * (64 * RW);
    if (!force && flag[FL_PASS2] == 0) {
        if (!scale) return;
        const double nrm = scal[SL_NA];
        if (!(nrm > 1e-300)) return;
        const R sc = (R)(1.0 / nrm);
        if (t < 64 * RW) {
            const i64 r = rb + t;
            if (r < n) {
                Cx<R> v = x[r];
                v.re *= sc;
                v.im *= sc;
                x[r] = v;
            }
        }
        return;
    }
    R hh = 0;
    for (int j = t; j < k; j += TPB) {
        const Cx<R> hj = h2[j];
        sh[j] = hj;
        hh += hj.re * hj.re + hj.im * hj.im;
    }
    hh = block_sum(hh, lds4);
    if (t == 0) {
        double sc = 1.0;
        int resc = 0;
        if (scale) {
            const double na = scal[SL_NA];
            const double e = na * na - (double)hh;
            if (na > 1e-300 && e >= trust * na * na) sc = 1.0 / sqrt(e);
            else resc = 1;
        }
        lsc = sc;
        if (vb.bx == 0) flag[FL_RESCALE] = resc;
    }
    __syncthreads();
    const R sc = (R)lsc;
    Cx<R> a[RW];
#pragma unroll
    for (int e = 0; e < RW; e++) a[e] = cx<R>(0, 0);
    if (rb + 64 * RW <= n) {
#pragma unroll 8
        for (int j = w; j < k; j += 4) {
            const Cx<R> hj = sh[j];
            const Cx<R> *q = Q + (i64)j * ldq + rb + lane;
#pragma unroll
            for (int e = 0; e < RW; e++) cfma(a[e], q[e * 64], hj);
        }
    } else {
        for (int j = w; j < k; j += 4) {
            const Cx<R> hj = sh[j];
            const Cx<R> *q = Q + (i64)j * ldq;
#pragma unroll
            for (int e = 0; e < RW; e++) {
                const i64 r = rb + lane + e * 64;
                if (r < n) cfma(a[e], q[r], hj);
            }
        }
    }
#pragma unroll
    for (int e = 0; e < RW; e++) red[w][lane + e * 64] = a[e];
    __syncthreads();
    R s = 0;
    if (t < 64 * RW) {
        const i64 r = rb + t;
        if (r < n) {
            Cx<R> v = x[r] - ((red[0][t] + red[1][t]) + (red[2][t] + red[3][t]));
            s = v.re * v.re + v.im * v.im;
            v.re *= sc;
            v.im *= sc;
            x[r] = v;
        }
    }
    if (fin.mode < 0) return;
    s = block_sum(s, lds4);
    if (t == 0) pnorm[vb.bx] = s;
    if (!fused || !last_block_of(cnt, vb.gx, &lflag)) return;
    finish_norm<R>(pnorm, vb.gx, fin, lds4);
    if (t == 0) *cnt = 0;
}
template <typename R, int RW>
__global__ __launch_bounds__(TPB) void k_cgs_final(i64 n, int k, const Cx<R> *__restrict__ Q, i64 ldq,
                                                   const Cx<R> *__restrict__ h2, Cx<R> *__restrict__ x, int scale,
                                                   int force, double trust, const double *scal, int *flag,
                                                   R *pnorm, int *cnt, NormFin fin, int fused)
{
    k_cgs_final_b<R, RW>(VB_HW, n, k, Q, ldq, h2, x, scale, force, trust, scal, flag, pnorm, cnt, fin, fused);
}

// y = Q[:, :k] c with the k <= 128 coefficients passed by value in the
// kernel arguments (no upload before the launch); columns split over the
// four waves as in k_update_split.
struct Coef128 { Cx<double> v[128]; };

template <int RW>
__global__ __launch_bounds__(TPB) void k_gemv_n_karg(i64 n, int k, const Cx<double> *__restrict__ Q, i64 ldq,
                                                     Coef128 c, Cx<double> *__restrict__ y)
{
    __shared__ Cx<double> red[4][64 * RW];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const i64 rb = (i64)blockIdx.x * (64 * RW);
    Cx<double> a[RW];
#pragma unroll
    for (int e = 0; e < RW; e++) a[e] = cx(0.0, 0.0);
    if (rb + 64 * RW <= n) {
#pragma unroll 8
        for (int j = w; j < k; j += 4) {
            const Cx<double> hj = c.v[j];
            const Cx<double> *q = Q + (i64)j * ldq + rb + lane;
#pragma unroll
            for (int e = 0; e < RW; e++) cfma(a[e], q[e * 64], hj);
        }
    } else {
        for (int j = w; j < k; j += 4) {
            const Cx<double> hj = c.v[j];
            const Cx<double> *q = Q + (i64)j * ldq;
#pragma unroll
            for (int e = 0; e < RW; e++) {
                const i64 r = rb + lane + e * 64;
                if (r < n) cfma(a[e], q[r], hj);
            }
        }
    }
#pragma unroll
    for (int e = 0; e < RW; e++) red[w][lane + e * 64] = a[e];
    __syncthreads();
    if (t < 64 * RW) {
        const i64 r = rb + t;
        if (r < n) y[r] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
    }
}

// separate finish of a norm (one workgroup)
template <typename R>
__device__ __forceinline__ void k_norm_fin_b(const VB vb, const R *pnorm, int nblk, NormFin fin,
                                             const int *__restrict__ cond)
{
    if (cond && cond[0] == 0) return;
    __shared__ R lds4[4];
    finish_norm<R>(pnorm, nblk, fin, lds4);
}
template <typename R>
__global__ __launch_bounds__(TPB) void k_norm_fin(const R *pnorm, int nblk, NormFin fin, const int *__restrict__ cond)
{
    k_norm_fin_b<R>(VB_HW, pnorm, nblk, fin, cond);
}

// ||x|| -> scal[slot] in one launch.
template <typename R, int E>
__device__ __forceinline__ void k_norm_b(const VB vb, i64 n, const Cx<R> *__restrict__ x, R *pnorm, int *cnt,
                                         NormFin fin,
                                              int fused)
{
    __shared__ R lds4[4];
    __shared__ int lflag;
    const i64 rb = (i64)vb.bx * (TPB * E);
    R s = 0;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const i64 r = rb + threadIdx.x + (i64)e * TPB;
        if (r < n) { const Cx<R> v = x[r]; s += v.re * v.re + v.im * v.im; }
    }
    s = block_sum(s, lds4);
    if (threadIdx.x == 0) pnorm[vb.bx] = s;
    if (!fused || !last_block_of(cnt, vb.gx, &lflag)) return;
    finish_norm<R>(pnorm, vb.gx, fin, lds4);
    if (threadIdx.x == 0) *cnt = 0;
}
template <typename R, int E>
__global__ __launch_bounds__(TPB) void k_norm(i64 n, const Cx<R> *__restrict__ x, R *pnorm, int *cnt, NormFin fin,
                                              int fused)
{
    k_norm_b<R, E>(VB_HW, n, x, pnorm, cnt, fin, fused);
}

template <typename R>
__device__ __forceinline__ void k_scale_inv_b(const VB vb, i64 n, Cx<R> *x, const double *scal, int slot, int mode,
                                                   const int *__restrict__ cond, Cx<R> *__restrict__ y)
{
    if (cond && cond[0] == 0) return;
    const double nrm = scal[slot];
    if (mode == 1 && !(nrm > 1e-300)) {
        if (y)
            for (i64 i = (i64)vb.bx * TPB + threadIdx.x; i < n; i += (i64)vb.gx * TPB) y[i] = x[i];
        return;
    }
    const R s = (R)(1.0 / nrm);
    for (i64 i = (i64)vb.bx * TPB + threadIdx.x; i < n; i += (i64)vb.gx * TPB) {
        Cx<R> v = x[i];
        v.re *= s; v.im *= s;
        x[i] = v;
        if (y) y[i] = v;
    }
}
template <typename R>
__global__ __launch_bounds__(TPB) void k_scale_inv(i64 n, Cx<R> *x, const double *scal, int slot, int mode,
                                                   const int *__restrict__ cond, Cx<R> *__restrict__ y)
{
    k_scale_inv_b<R>(VB_HW, n, x, scal, slot, mode, cond, y);
}

template <typename R>
__device__ __forceinline__ void k_scale_b(const VB vb, i64 n, Cx<R> *x, R are, R aim)
{
    const Cx<R> a = cx(are, aim);
    for (i64 i = (i64)vb.bx * TPB + threadIdx.x; i < n; i += (i64)vb.gx * TPB) x[i] = a * x[i];
}
template <typename R>
__global__ __launch_bounds__(TPB) void k_scale(i64 n, Cx<R> *x, R are, R aim)
{
    k_scale_b<R>(VB_HW, n, x, are, aim);
}

// out = a x + b y; with co non-NULL the coefficients are read from device
// memory instead, a = (co[0], co[1]), b = (co[2], co[3]) (doubles, rounded
// to R as host arguments are), so a captured graph replays with new ones.
template <typename R, int E>
__device__ __forceinline__ void k_lincomb_b(const VB vb, i64 n, R are, R aim, const Cx<R> *x, R bre, R bim,
                                                 const Cx<R> *y, Cx<R> *out, R *pnorm, int *cnt, NormFin fin,
                                                 int fused, const double *__restrict__ co)
{
    __shared__ R lds4[4];
    __shared__ int lflag;
    if (co) { are = (R)co[0]; aim = (R)co[1]; bre = (R)co[2]; bim = (R)co[3]; }
    const Cx<R> a = cx(are, aim), b = cx(bre, bim);
    const i64 rb = (i64)vb.bx * (TPB * E);
    R s = 0;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const i64 r = rb + threadIdx.x + (i64)e * TPB;
        if (r < n) {
            const Cx<R> v = a * x[r] + b * y[r];
            out[r] = v;
            s += v.re * v.re + v.im * v.im;
        }
    }
    if (fin.mode < 0) return;
    s = block_sum(s, lds4);
    if (threadIdx.x == 0) pnorm[vb.bx] = s;
    if (!fused || !last_block_of(cnt, vb.gx, &lflag)) return;
    finish_norm<R>(pnorm, vb.gx, fin, lds4);
    if (threadIdx.x == 0) *cnt = 0;
}
template <typename R, int E>
__global__ __launch_bounds__(TPB) void k_lincomb(i64 n, R are, R aim, const Cx<R> *x, R bre, R bim,
                                                 const Cx<R> *y, Cx<R> *out, R *pnorm, int *cnt, NormFin fin,
                                                 int fused, const double *__restrict__ co)
{
    k_lincomb_b<R, E>(VB_HW, n, are, aim, x, bre, bim, y, out, pnorm, cnt, fin, fused, co);
}

// Sleijpen-van der Vorst step z += c s, with c = -(u^H z)/(u^H s) when
// |u^H s| > 1e-300, else 0 (reference src/jd_core.c:966-971), computed by
// every thread from the two dots d = (u^H z, u^H s).
__global__ __launch_bounds__(TPB) void k_axpy_svdv(i64 n, const Cx<double> *d, const Cx<double> *s, Cx<double> *z)
{
    const Cx<double> uz = d[0], us = d[1];
    const double a = hypot(us.re, us.im);
    Cx<double> c = cx(0.0, 0.0);
    if (a > 1e-300) {
        const double den = us.re * us.re + us.im * us.im;
        c = cx(-(uz.re * us.re + uz.im * us.im) / den, -(uz.im * us.re - uz.re * us.im) / den);
    }
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < n; i += (i64)gridDim.x * TPB) {
        Cx<double> v = z[i];
        cfma(v, s[i], c);
        z[i] = v;
    }
}

// k_axpy_svdv reducing the two dots itself from the projection's per-tile
// partials (part[(tile 2 + v)], v = 0: u^H z, 1: u^H s; ntiles tiles), in
// proj_finish's order (16 lanes, tile = lane mod 16, then a 16-lane xor
// butterfly), so every block holds the finish's values bit for bit and the
// finish needs no launch of its own; block 0 also stores them in d.
__global__ __launch_bounds__(TPB) void k_axpy_svdv_part(i64 n, const Cx<double> *__restrict__ part, int ntiles,
                                                        Cx<double> *d, const Cx<double> *s, Cx<double> *z)
{
    __shared__ Cx<double> dots[2];
    const int t = threadIdx.x;
    if (t < 32) {
        const int vi = t >> 4, l = t & 15;
        double sr = 0, si = 0;
        for (int tt = l; tt < ntiles; tt += 16) {
            const Cx<double> p = part[(i64)tt * 2 + vi];
            sr += p.re;
            si += p.im;
        }
#pragma unroll
        for (int m = 8; m >= 1; m >>= 1) {
            sr += __shfl_xor(sr, m, 64);
            si += __shfl_xor(si, m, 64);
        }
        if (l == 0) {
            dots[vi] = cx(sr, si);
            if (blockIdx.x == 0) d[vi] = cx(sr, si);
        }
    }
    __syncthreads();
    const Cx<double> uz = dots[0], us = dots[1];
    const double a = hypot(us.re, us.im);
    Cx<double> c = cx(0.0, 0.0);
    if (a > 1e-300) {
        const double den = us.re * us.re + us.im * us.im;
        c = cx(-(uz.re * us.re + uz.im * us.im) / den, -(uz.im * us.re - uz.re * us.im) / den);
    }
    for (i64 i = (i64)blockIdx.x * TPB + t; i < n; i += (i64)gridDim.x * TPB) {
        Cx<double> v = z[i];
        cfma(v, s[i], c);
        z[i] = v;
    }
}

// In-place basis rotation X[:, :keep] = X[:, :k] Z (Zr row-major k x keep).
// A block stages RT rows x k columns of X in LDS, so every input of its rows
// is read before any output is written; outputs go back through LDS for
// column-contiguous stores.
template <int RT>
__global__ __launch_bounds__(TPB) void k_rotate(i64 n, int k, int keep, Cx<double> *X, i64 ldx,
                                                const Cx<double> *__restrict__ Zr)
{
    extern __shared__ char smem_raw[];
    Cx<double> *tin = reinterpret_cast<Cx<double> *>(smem_raw);     // k x RT
    Cx<double> *tout = tin + (i64)k * RT;                            // keep x RT
    const int t = threadIdx.x;
    const i64 rb = (i64)blockIdx.x * RT;
    for (int idx = t; idx < RT * k; idx += TPB) {
        const int j = idx / RT, rr = idx % RT;
        const i64 r = rb + rr;
        tin[idx] = r < n ? X[(i64)j * ldx + r] : cx(0.0, 0.0);
    }
    __syncthreads();
    constexpr int RPT = RT / 4;            // rows per thread
    const int cl = t & 63, rg = t >> 6;    // column lane, row group
    for (int cb = 0; cb < keep; cb += 64) {
        const int c = cb + cl;
        Cx<double> a[RPT];
#pragma unroll
        for (int q = 0; q < RPT; q++) a[q] = cx(0.0, 0.0);
        if (c < keep) {
            for (int j = 0; j < k; j++) {
                const Cx<double> z = Zr[(i64)j * keep + c];
#pragma unroll
                for (int q = 0; q < RPT; q++) cfma(a[q], tin[j * RT + rg * RPT + q], z);
            }
#pragma unroll
            for (int q = 0; q < RPT; q++) tout[c * RT + rg * RPT + q] = a[q];
        }
    }
    __syncthreads();
    for (int idx = t; idx < RT * keep; idx += TPB) {
        const int c = idx / RT, rr = idx % RT;
        const i64 r = rb + rr;
        if (r < n) X[(i64)c * ldx + r] = tout[idx];
    }
}

// The same rotation on the FP64 matrix cores: a block stages RT rows x k
// columns of X in LDS (every input of its rows is read before any output is
// written), and the RT x keep output = (RT x k) Z is cut into 16 x 16 complex
// blocks, wave w taking blocks w, w + 4, ...; per 4-column step of the inner
// dimension four v_mfma_f64_16x16x4f64 (Re += Xr Zr - Xi Zi, Im += Xr Zi +
// Xi Zr; operand layout as in k_gram).  Z (Zr row-major, k x keep) is read
// from global memory (it is small and shared by every block).  The output goes
// back through the same LDS for column-contiguous stores.  RT * ceil4(k) * 16
// bytes of LDS; at most 16 output blocks (2 RT / 16 * ceil(keep / 16) <= 16).
template <int RT>
__global__ __launch_bounds__(TPB) void k_rotate_mfma(i64 n, int k, int keep, Cx<double> *X, i64 ldx,
                                                     const Cx<double> *__restrict__ Zr)
{
    extern __shared__ char smem_raw[];
    Cx<double> *tl = reinterpret_cast<Cx<double> *>(smem_raw);     // column j, row r at j RT + r
    constexpr int MT = RT / 16;                                     // row blocks
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int ml = lane & 15, kq = lane >> 4;
    const int k4 = (k + 3) & ~3;
    const i64 rb = (i64)blockIdx.x * RT;
    for (int idx = t; idx < RT * k4; idx += TPB) {
        const int j = idx / RT, rr = idx % RT;
        const i64 r = rb + rr;
        const bool ok = j < k && r < n;
        Cx<double> v = X[ok ? (i64)j * ldx + r : 0];
        if (!ok) v = cx(0.0, 0.0);
        tl[idx] = v;
    }
    __syncthreads();
    const int nblk = MT * ((keep + 15) / 16);
    v4d acr[4], aci[4];
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
        acr[s4] = v4d{0, 0, 0, 0};
        aci[s4] = v4d{0, 0, 0, 0};
    }
    for (int kk = 0; kk < k4; kk += 4) {
        const int kr = kk + kq;
#pragma unroll
        for (int s4 = 0; s4 < 4; s4++) {
            const int b = w + 4 * s4;
            if (b < nblk) {                                         // uniform over the wave
                const int mt = b % MT, nt = b / MT;
                const Cx<double> a = tl[kr * RT + mt * 16 + ml];
                const int col = nt * 16 + ml;
                const bool ok = col < keep && kr < k;
                Cx<double> z = Zr[ok ? (i64)kr * keep + col : 0];
                if (!ok) z = cx(0.0, 0.0);
                acr[s4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, z.re, acr[s4], 0, 0, 0);
                acr[s4] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, z.im, acr[s4], 0, 0, 0);
                aci[s4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, z.im, aci[s4], 0, 0, 0);
                aci[s4] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, z.re, aci[s4], 0, 0, 0);
            }
        }
    }
    __syncthreads();                                                // every read of the inputs done
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
        const int b = w + 4 * s4;
        if (b < nblk) {
            const int mt = b % MT, nt = b / MT, col = nt * 16 + ml;
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (col < keep) tl[col * RT + mt * 16 + 4 * e + kq] = cx(acr[s4][e], aci[s4][e]);
        }
    }
    __syncthreads();
    for (int idx = t; idx < RT * keep; idx += TPB) {
        const int c = idx / RT, rr = idx % RT;
        const i64 r = rb + rr;
        if (r < n) X[(i64)c * ldx + r] = tl[idx];
    }
}

// Streaming form of the rotation on the FP64 matrix cores: a block owns 64
// rows and all keep output columns; wave w owns the output's rows
// 16w .. 16w+15 against every 16-column block (NT of them), so its A operand
// is loaded once per 4-column step and reused NT times.  The k input columns
// (and the matching rows of Z) stream through two LDS buffers in chunks of
// KC: the next chunk's global loads are in flight while the MFMAs consume
// the current one.  Outputs stay in registers until every input column of
// the block's rows has been read, then go straight to X (in place); each
// 16 x 16 block is stored as four consecutive rows per column (64 bytes).
// LDS: 2 (64 KC + KC 16 NT) * 16 bytes, independent of k.
template <int NT, int KC>
__global__ __launch_bounds__(TPB) void k_rotate_mfma_s(i64 n, int k, int keep, Cx<double> *X, i64 ldx,
                                                       const Cx<double> *__restrict__ Zr)
{
    constexpr int RT = 64, KP = 16 * NT;                 // rows per block, padded output columns
    constexpr int XE = RT * KC / TPB, ZE = KC * KP / TPB;  // staged elements per thread
    static_assert(XE >= 1 && ZE >= 1 && RT * KC % TPB == 0 && KC * KP % TPB == 0, "tile shape");
    __shared__ Cx<double> xs[2][KC * RT];                // column j of the chunk, row r at j RT + r
    __shared__ Cx<double> zs[2][KC * KP];                // row j of the chunk's Z, column c at j KP + c
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int ml = lane & 15, kq = lane >> 4;
    const i64 rb = (i64)blockIdx.x * RT;
    const int nch = (k + KC - 1) / KC;
    Cx<double> px[XE], pz[ZE];
    auto load = [&](int c) {
        const int j0 = c * KC;
#pragma unroll
        for (int e = 0; e < XE; e++) {
            const int idx = t + e * TPB, jj = idx / RT, rr = idx % RT;
            const int j = j0 + jj;
            const i64 r = rb + rr;
            const bool ok = j < k && r < n;
            Cx<double> v = X[ok ? (i64)j * ldx + r : 0];
            px[e] = ok ? v : cx(0.0, 0.0);
        }
#pragma unroll
        for (int e = 0; e < ZE; e++) {
            const int idx = t + e * TPB, jj = idx / KP, cc = idx % KP;
            const int j = j0 + jj;
            const bool ok = j < k && cc < keep;
            Cx<double> v = Zr[ok ? (i64)j * keep + cc : 0];
            pz[e] = ok ? v : cx(0.0, 0.0);
        }
    };
    auto stage = [&](int b) {
#pragma unroll
        for (int e = 0; e < XE; e++) xs[b][t + e * TPB] = px[e];
#pragma unroll
        for (int e = 0; e < ZE; e++) zs[b][t + e * TPB] = pz[e];
    };
    v4d acr[NT], aci[NT];
#pragma unroll
    for (int b = 0; b < NT; b++) {
        acr[b] = v4d{0, 0, 0, 0};
        aci[b] = v4d{0, 0, 0, 0};
    }
    load(0);
    stage(0);
    __syncthreads();
    for (int c = 0; c < nch; c++) {
        const int cur = c & 1;
        if (c + 1 < nch) load(c + 1);                    // in flight during the MFMAs below
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            const Cx<double> a = xs[cur][(kk + kq) * RT + w * 16 + ml];
#pragma unroll
            for (int b = 0; b < NT; b++) {
                const Cx<double> z = zs[cur][(kk + kq) * KP + b * 16 + ml];
                acr[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, z.re, acr[b], 0, 0, 0);
                acr[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, z.im, acr[b], 0, 0, 0);
                aci[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, z.im, aci[b], 0, 0, 0);
                aci[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, z.re, aci[b], 0, 0, 0);
            }
        }
        // the other buffer's last readers finished before the previous barrier
        if (c + 1 < nch) stage(cur ^ 1);
        __syncthreads();
    }
    // D(m, n) in lane n + 16 (m % 4), element m / 4: rows 16 w + 4 e + kq
#pragma unroll
    for (int b = 0; b < NT; b++) {
        const int col = b * 16 + ml;
        if (col < keep)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const i64 r = rb + w * 16 + 4 * e + kq;
                if (r < n) X[(i64)col * ldx + r] = cx(acr[b][e], aci[b][e]);
            }
    }
}

// Gram block G = X^H Y of two tall-skinny complex matrices (X n x p, Y n x q)
// on the FP64 matrix cores.  A block owns a 64 x 64 tile of G and a chunk of
// rows; the chunk is staged through LDS in slabs of GR_SR rows (coalesced
// column reads, zeros outside X and Y).  Wave w computes the tile's rows
// 16w .. 16w+15 against its four 16-column groups with
// v_mfma_f64_16x16x4f64: per 4-row step, Re += Xr^T Yr + Xi^T Yi and
// Im += Xr^T Yi - Xi^T Yr (four MFMAs per 16 x 16 complex block).
// Operand layout of the 16x16x4 f64 MFMA (checked on gfx950): A(m, k) in
// lane m + 16 k, B(k, n) in lane n + 16 k, and D(m, n) in lane n + 16 (m % 4),
// element m / 4.  Partials per chunk go to part[chunk][i + j p];
// k_gram_finish adds the chunks in a fixed order.
static constexpr int GR_T = 64, GR_SR = 32, GR_LD = GR_SR + 1;   // +1: LDS banks

__global__ __launch_bounds__(TPB) void k_gram(i64 n, i64 chunk, int p, int q, const Cx<double> *__restrict__ X,
                                              i64 ldx, const Cx<double> *__restrict__ Y, i64 ldy,
                                              Cx<double> *__restrict__ part)
{
    __shared__ Cx<double> xs[GR_T * GR_LD], ys[GR_T * GR_LD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int tiles_q = (q + GR_T - 1) / GR_T;
    const int i0 = (blockIdx.x / tiles_q) * GR_T, j0 = (blockIdx.x % tiles_q) * GR_T;
    const i64 rb = (i64)blockIdx.y * chunk;
    const i64 re = rb + chunk < n ? rb + chunk : n;
    const int ml = lane & 15, kq = lane >> 4;
    v4d accr[4], acci[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        accr[j] = v4d{0, 0, 0, 0};
        acci[j] = v4d{0, 0, 0, 0};
    }
    // The slab's loads are issued into registers one slab ahead, branch-free
    // (an element outside X or Y loads X[0] / Y[0] and is zeroed), so HBM
    // latency overlaps the previous slab's MFMAs.
    constexpr int E = GR_T * GR_SR / TPB;          // elements per thread and matrix
    Cx<double> xv[E], yv[E];
    auto load_slab = [&](i64 r0) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int idx = t + e * TPB, rr = idx % GR_SR, cc = idx / GR_SR;
            const i64 r = r0 + rr;
            const bool okx = r < re && i0 + cc < p, oky = r < re && j0 + cc < q;
            xv[e] = X[okx ? (i64)(i0 + cc) * ldx + r : 0];
            yv[e] = Y[oky ? (i64)(j0 + cc) * ldy + r : 0];
            if (!okx) xv[e] = cx(0.0, 0.0);
            if (!oky) yv[e] = cx(0.0, 0.0);
        }
    };
    if (rb < re) load_slab(rb);
    for (i64 r0 = rb; r0 < re; r0 += GR_SR) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int idx = t + e * TPB, rr = idx % GR_SR, cc = idx / GR_SR;
            xs[cc * GR_LD + rr] = xv[e];
            ys[cc * GR_LD + rr] = yv[e];
        }
        __syncthreads();
        if (r0 + GR_SR < re) load_slab(r0 + GR_SR);
#pragma unroll 2
        for (int kk = 0; kk < GR_SR; kk += 4) {
            const Cx<double> a = xs[(16 * w + ml) * GR_LD + kk + kq];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const Cx<double> bv = ys[(16 * j + ml) * GR_LD + kk + kq];
                accr[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, bv.re, accr[j], 0, 0, 0);
                accr[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, bv.im, accr[j], 0, 0, 0);
                acci[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, bv.im, acci[j], 0, 0, 0);
                acci[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, bv.re, acci[j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    Cx<double> *pc = part + (i64)blockIdx.y * p * q;
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int i = i0 + 16 * w + 4 * e + kq, jj = j0 + 16 * j + ml;
            if (i < p && jj < q) pc[(i64)jj * p + i] = cx(accr[j][e], acci[j][e]);
        }
}

// G[i, j] = sum over chunks of the partials: 16 lanes per entry, lane l adds
// chunks l, l + 16, ... in order, then a fixed xor tree over the 16 lanes.
__global__ __launch_bounds__(TPB) void k_gram_finish(int nch, int p, int q, const Cx<double> *__restrict__ part,
                                                     Cx<double> *G, i64 ldg)
{
    const i64 pq = (i64)p * q;
    const i64 e = ((i64)blockIdx.x * TPB + threadIdx.x) >> 4;
    const int l = threadIdx.x & 15;
    double sr = 0, si = 0;
    if (e < pq)
        for (int c = l; c < nch; c += 16) {
            const Cx<double> v = part[(i64)c * pq + e];
            sr += v.re;
            si += v.im;
        }
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) {
        sr += __shfl_xor(sr, m, 64);
        si += __shfl_xor(si, m, 64);
    }
    if (l == 0 && e < pq) {
        const int i = (int)(e % p), j = (int)(e / p);
        G[(i64)j * ldg + i] = cx(sr, si);
    }
}

__device__ __forceinline__ void k_z2c_b(const VB vb, i64 n, const Cx<double> *z, Cx<float> *c)
{
    for (i64 i = (i64)vb.bx * TPB + threadIdx.x; i < n; i += (i64)vb.gx * TPB)
        c[i] = cx((float)z[i].re, (float)z[i].im);
}
__global__ __launch_bounds__(TPB) void k_z2c(i64 n, const Cx<double> *z, Cx<float> *c)
{
    k_z2c_b(VB_HW, n, z, c);
}

__device__ __forceinline__ void k_c2z_b(const VB vb, i64 n, const Cx<float> *c, Cx<double> *z)
{
    for (i64 i = (i64)vb.bx * TPB + threadIdx.x; i < n; i += (i64)vb.gx * TPB)
        z[i] = cx((double)c[i].re, (double)c[i].im);
}
__global__ __launch_bounds__(TPB) void k_c2z(i64 n, const Cx<float> *c, Cx<double> *z)
{
    k_c2z_b(VB_HW, n, c, z);
}

// ---------------------------------------------------------------------------
// host launchers
// Row tiles.  A block of the norm / lincomb kernels, and of the update for
// k < 16, owns TPB * E rows, E in {1, 2, 4}, chosen per launch from the
// shape: the largest tile that still gives g_blocks_target blocks.  From
// k = 16 the update splits the columns over the block's waves (128 rows per
// block).  A block of the projection kernel streams a chunk of rows (a
// multiple of BBJD_PROJ_MIN_ROWS) sized so that the launch has about
// g_proj_blocks blocks: long chunks amortise the block's cross-lane
// reduction, enough blocks keep every CU streaming.  The defaults come from
// sweeps of n in [2e4, 1e6], k in [8, 128] on MI355X; all can be forced for
// tuning.
// Partial-sum buffers are sized for the smallest tiles, BBJD_TILE_MIN_ROWS
// (norm partials, one per block) and BBJD_PROJ_MIN_ROWS (projection chunks).
static int g_upd_rows = 0;                      // 0: choose from the shape
static int g_blocks_target = 2048;
static int g_proj_blocks = 192;

// Per-shape tile table (filled by the autotuner, bbjd_amd/autotune.py):
// class (i, j) holds the shapes with n <= n_upper[i] and k <= k_upper[j]
// (the last class of each takes everything larger); entries are the
// projection's blocks per launch and the update's rows per block (0: the
// shape rule).  Empty: the global defaults above.  Set before solves run.
static constexpr int TT_MAX = 8;
static int g_tt_nn = 0, g_tt_nk = 0;
static i64 g_tt_n[TT_MAX];
static int g_tt_k[TT_MAX], g_tt_proj[TT_MAX * TT_MAX], g_tt_upd[TT_MAX * TT_MAX];

static inline int tt_class(i64 n, int k)
{
    int i = 0, j = 0;
    while (i < g_tt_nn - 1 && n > g_tt_n[i]) i++;
    while (j < g_tt_nk - 1 && k > g_tt_k[j]) j++;
    return i * TT_MAX + j;
}

static inline int ntiles(i64 n, int rows) { return (int)((n + rows - 1) / rows); }
static inline int pick_rows(i64 n, int col_tiles)
{
    for (int rows = 4 * TPB; rows > TPB; rows >>= 1)
        if ((i64)ntiles(n, rows) * col_tiles >= g_blocks_target) return rows;
    return TPB;
}
static inline int upd_rows(i64 n, int k)
{
    if (g_upd_rows) return g_upd_rows;
    if (g_tt_nn > 0) {
        const int r = g_tt_upd[tt_class(n, k)];
        if (r) return r;
    }
    return k >= 16 ? 128 : pick_rows(n, 1);     // measured: column split wins from k = 16
}
static inline int vec_rows(i64 n) { return g_upd_rows >= TPB ? g_upd_rows : pick_rows(n, 1); }
static inline i64 proj_chunk(i64 n, int col_tiles, int k)
{
    const int blocks = g_tt_nn > 0 ? g_tt_proj[tt_class(n, k)] : g_proj_blocks;
    i64 chunks = (blocks + col_tiles - 1) / col_tiles;
    if (chunks < 1) chunks = 1;
    i64 chunk = (n + chunks - 1) / chunks;
    chunk = (chunk + BBJD_PROJ_MIN_ROWS - 1) / BBJD_PROJ_MIN_ROWS * BBJD_PROJ_MIN_ROWS;
    return chunk > 0 ? chunk : BBJD_PROJ_MIN_ROWS;
}
static inline int grid_stride_blocks(i64 n)
{
    i64 b = (n + TPB - 1) / TPB;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}
static inline int chk(void) { return hipGetLastError() == hipSuccess ? 0 : -1; }

extern "C" int64_t bbjd_red_part_entries(int64_t n, int maxk)
{
    // up to three vectors per projection (the test-space expansion), and at
    // least 128 row chunks of a maxk x maxk Gram matrix (the restart)
    const int64_t proj = (int64_t)ntiles(n, BBJD_PROJ_MIN_ROWS) * 3 * (maxk + 1);
    const int64_t gram = (int64_t)128 * maxk * maxk;
    return (proj > gram ? proj : gram) + 64;
}

// Blocks up to which the finishing step runs in the producing kernel's last
// block; larger grids finish in a second small kernel instead.  Each fused
// block pays an agent-scope release (an L2 write-back), which costs more
// than a launch beyond a few blocks -- and far more when other streams' data
// sits dirty in the L2: on the 16-stream serving workload a threshold of 16
// served 160 pairs/s, 48 served 153 and fusing every grid 108.  Tunable.
static int g_fuse_max_blocks = 16;
// SvdV: the axpy launch reduces the two dots' partials itself (1) or a
// finish launch does (0)
static int g_svdv_fused = 1;

static inline NormFin fin_none()
{
    NormFin f;
    f.scal = nullptr; f.flag = nullptr; f.mode = -1; f.slot = 0; f.k = 0;
    f.hsum = nullptr; f.h = nullptr; f.h2 = nullptr; f.x = nullptr; f.n = 0;
    return f;
}
static inline NormFin fin_norm(bbjd_red *rd, int mode, int slot)
{
    NormFin f = fin_none();
    f.scal = rd->scal; f.flag = rd->flag; f.mode = mode; f.slot = slot;
    return f;
}

#define BBJD_E_SWITCH(rows, CALL) \
    switch (rows) {                  \
    case 4 * TPB: CALL(4); break;    \
    case 2 * TPB: CALL(2); break;    \
    default: CALL(1); break;         \
    }

// One projection pass in one launch (+ a finishing launch for large grids):
// out0 (and out1) <- Q^H x0 (x1); with norm_slot >= 0 also scal[norm_slot]
// <- ||x0||.
template <typename R, int NV>
static int proj_pass_nv(hipStream_t s, i64 n, int k, const void *Q, i64 ldq, const void *x0, const void *x1,
                        const void *x2, void *out0, void *out1, void *out2, bbjd_red *rd, int norm_slot,
                        const int *cond, Cx<R> ca = Cx<R>{0, 0}, Cx<R> cb = Cx<R>{0, 0},
                        int *finish_left = nullptr)
{
    constexpr int C = NV == 1 ? 16 : 8;
    const int ctiles = k > 0 ? (k + C - 1) / C : 1;
    const i64 chunk = proj_chunk(n, ctiles, k);
    const int nt = (int)((n + chunk - 1) / chunk);
    const int ldp = k > 0 ? k : 1;
    const int want_norm = norm_slot >= 0;
    if ((i64)nt * NV * ldp > rd->part_cap || k > BBJD_MAXK) return -2;
    if (k == 0 && !want_norm) return 0;
    const int fused = nt <= g_fuse_max_blocks;
    dim3 g(ctiles, nt);
    bool timed = NV == 1 && g_kt_on && sizeof(R) == 8 && cond == nullptr;
    if (timed) {
        // a stream under graph capture records no event-carrying dispatch
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) timed = false;
    }
    hipEvent_t ea = nullptr, eb = nullptr;
    if (timed) {
        // a bounded record: each recorded event keeps its dispatch (and what
        // that dispatch references) alive until harvested
        std::lock_guard<std::mutex> lk(g_kt_mu);
        if (g_kt.size() >= kKtMaxRecords) timed = false;
        else { ea = kt_event(); eb = kt_event(); }
    }
    if (timed) {
        // events carried by the dispatch itself: their interval is the
        // kernel's execution, as a kernel trace reports it
        hipExtLaunchKernelGGL((k_proj<R, NV, C, false>), g, dim3(TPB), 0, s, ea, eb, 0, (i64)n, chunk, k,
                              (const Cx<R> *)Q, (i64)ldq, (const Cx<R> *)x0, (const Cx<R> *)x1, (const Cx<R> *)x2,
                              ca, cb, (Cx<R> *)rd->part, ldp, (R *)rd->pnorm, want_norm, cond, rd->cnt, (Cx<R> *)out0,
                              (Cx<R> *)out1, (Cx<R> *)out2, rd->scal, norm_slot, fused);
        const double bytes = (double)sizeof(Cx<R>) * (double)n * ((double)k + 1.0);
        std::lock_guard<std::mutex> lk(g_kt_mu);
        g_kt.push_back({ea, eb, bytes});
    } else if (cond)
        bbjd::gl<&k_proj_b<R, NV, C, true>, TPB>(k_proj<R, NV, C, true>, g, 0, s, (i64)n, chunk, k, (const Cx<R> *)Q,
                           (i64)ldq, (const Cx<R> *)x0, (const Cx<R> *)x1, (const Cx<R> *)x2, ca, cb,
                            (Cx<R> *)rd->part, ldp,
                           (R *)rd->pnorm, want_norm, cond, rd->cnt, (Cx<R> *)out0, (Cx<R> *)out1, (Cx<R> *)out2,
                           rd->scal, norm_slot, fused);
    else
        bbjd::gl<&k_proj_b<R, NV, C, false>, TPB>(k_proj<R, NV, C, false>, g, 0, s, (i64)n, chunk, k, (const Cx<R> *)Q,
                           (i64)ldq, (const Cx<R> *)x0, (const Cx<R> *)x1, (const Cx<R> *)x2, ca, cb,
                            (Cx<R> *)rd->part, ldp,
                           (R *)rd->pnorm, want_norm, cond, rd->cnt, (Cx<R> *)out0, (Cx<R> *)out1, (Cx<R> *)out2,
                           rd->scal, norm_slot, fused);
    // finish_left: a caller that reduces the partials itself gets the tile
    // count instead of a finish launch (0 when the last block finished)
    if (finish_left) *finish_left = fused ? 0 : nt;
    if (!fused && !finish_left)
        bbjd::gl<&k_proj_finish_b<R, NV, C>, TPB>(k_proj_finish<R, NV, C>, dim3(g.x), 0, s, nt, k,
                                                  (const Cx<R> *)rd->part,
                           ldp, (const R *)rd->pnorm, want_norm, cond, (Cx<R> *)out0, (Cx<R> *)out1,
                           (Cx<R> *)out2, rd->scal, norm_slot);
    return chk();
}

template <typename R>
static int proj_pass(hipStream_t s, i64 n, int k, const void *Q, i64 ldq, const void *x0, const void *x1,
                     void *out0, void *out1, bbjd_red *rd, int norm_slot, const int *cond)
{
    return x1 ? proj_pass_nv<R, 2>(s, n, k, Q, ldq, x0, x1, nullptr, out0, out1, nullptr, rd, norm_slot, cond)
              : proj_pass_nv<R, 1>(s, n, k, Q, ldq, x0, nullptr, nullptr, out0, nullptr, nullptr, rd, norm_slot,
                                   cond);
}

template <typename R>
static int update_pass(hipStream_t s, i64 n, int k, const void *Q, i64 ldq, const void *h, R alpha, int beta_one,
                       void *x, bbjd_red *rd, NormFin fin, const int *cond)
{
    if (k > BBJD_MAXK) return -2;
    const int rows = upd_rows(n, k);
    const int nt = ntiles(n, rows);
    const size_t lds = (size_t)(k > 0 ? k : 1) * sizeof(Cx<R>);
    const int fused = nt <= g_fuse_max_blocks;
    if (rows < TPB) {
#define BBJD_LAUNCH_SPLIT(RW, CO)                                                                                  \
    bbjd::gl<&k_update_split_b<R, RW, CO>, TPB>(k_update_split<R, RW, CO>, dim3(nt), lds, s, (i64)n, k,              \
                                                (const Cx<R> *)Q,                                                    \
                       (i64)ldq, (const Cx<R> *)h, alpha, beta_one, (Cx<R> *)x, (R *)rd->pnorm, cond,                \
                       rd->cnt + BBJD_CNT_UPD, fin, fused)
        if (rows == 128) {
            if (cond) BBJD_LAUNCH_SPLIT(2, true);
            else BBJD_LAUNCH_SPLIT(2, false);
        } else {
            if (cond) BBJD_LAUNCH_SPLIT(1, true);
            else BBJD_LAUNCH_SPLIT(1, false);
        }
#undef BBJD_LAUNCH_SPLIT
        if (!fused && fin.mode >= 0)
            bbjd::gl<&k_norm_fin_b<R>, TPB>(k_norm_fin<R>, dim3(1), 0, s, (const R *)rd->pnorm, nt, fin, cond);
        return chk();
    }
#define BBJD_LAUNCH_UPD(E)                                                                                         \
    bbjd::gl<&k_update_b<R, E>, TPB>(k_update<R, E>, dim3(nt), lds, s, (i64)n, k, (const Cx<R> *)Q, (i64)ldq,         \
                       (const Cx<R> *)h, alpha, beta_one, (Cx<R> *)x, (R *)rd->pnorm, cond, rd->cnt + BBJD_CNT_UPD, \
                       fin, fused)
    BBJD_E_SWITCH(rows, BBJD_LAUNCH_UPD)
#undef BBJD_LAUNCH_UPD
    if (!fused && fin.mode >= 0)
        bbjd::gl<&k_norm_fin_b<R>, TPB>(k_norm_fin<R>, dim3(1), 0, s, (const R *)rd->pnorm, nt, fin, cond);
    return chk();
}

template <typename R>
static int norm_into(hipStream_t s, i64 n, const void *x, bbjd_red *rd, int slot)
{
    const int rows = vec_rows(n);
    const int nt = ntiles(n, rows);
    const int fused = nt <= g_fuse_max_blocks;
#define BBJD_LAUNCH_NORM(E)                                                                                        \
    bbjd::gl<&k_norm_b<R, E>, TPB>(k_norm<R, E>, dim3(nt), 0, s, (i64)n, (const Cx<R> *)x, (R *)rd->pnorm,         \
                       rd->cnt + BBJD_CNT_VEC, fin_norm(rd, 0, slot), fused)
    BBJD_E_SWITCH(rows, BBJD_LAUNCH_NORM)
#undef BBJD_LAUNCH_NORM
    if (!fused)
        bbjd::gl<&k_norm_fin_b<R>, TPB>(k_norm_fin<R>, dim3(1), 0, s, (const R *)rd->pnorm, nt,
                           fin_norm(rd, 0, slot), (const int *)nullptr);
    return chk();
}

// Blocks per Gram launch (row chunks x 64 x 64 tiles); tunable.
static int g_gram_blocks = 512;
// In-place rotation on the matrix cores (1) or on vector FP64 (0).
static int g_rotate_mfma = 1;
// Input columns per LDS chunk of the streaming rotation (0: per output
// width: 16 up to 32 output columns, else 8, which keeps more blocks
// resident; measured at n = 2e5, k = 128 -> 64: 287 us at 8, 299 at 16,
// 334 at 32); tunable.
static int g_rot_kc = 0;

// Fused CGS2 (k <= 128): on by default, switchable for tests and tuning
// (2: on, with the second pass's normalisation always by the computed norm).
static int g_cgs_fused = 1;

// columns per wave of k_cgs_mid for k columns; 0 beyond its register budget
static inline int cgs_kw(int k)
{
    return k <= 8 ? 2 : k <= 16 ? 4 : k <= 32 ? 8 : k <= 64 ? 16 : k <= 128 ? 32 : 0;
}

template <typename R>
static int scale_inv(hipStream_t s, i64 n, void *x, const double *scal, int slot, const int *cond)
{
    bbjd::gl<&k_scale_inv_b<R>, TPB>(k_scale_inv<R>, dim3(grid_stride_blocks(n)), 0, s, (i64)n, (Cx<R> *)x, scal,
                       slot, 1, cond, (Cx<R> *)nullptr);
    return chk();
}

// Middle launch of the fused CGS2 (+ its finish when the grid is large).
// Blocks walk enough 64-row tiles for at most g_mid_blocks blocks, and more
// when the partials (blocks x k) would not fit the scratch; the finish
// reduces one partial per block and column.  512 measured best on MI355X
// (n = 5e4, k = 32: 27.3 us for the fused projection against 33.4 with
// 1024; n = 2e5, k = 64: 105 against 112).  Tunable.
static int g_mid_blocks = 512;

template <typename R>
static int cgs_mid(hipStream_t s, i64 n, int k, const void *Q, i64 ldq, void *v, bbjd_red *rd, MidFin f,
                   const void *xa = nullptr, const void *xb = nullptr, Cx<R> ca = Cx<R>{0, 0},
                   Cx<R> cb = Cx<R>{0, 0})
{
    const i64 tiles64 = (n + 63) / 64;
    i64 per = (tiles64 + g_mid_blocks - 1) / g_mid_blocks;
    if (per < 1) per = 1;
    i64 nb = (tiles64 + per - 1) / per;
    while (nb * k > rd->part_cap && nb > 1) {
        per *= 2;
        nb = (tiles64 + per - 1) / per;
    }
    if (nb * k > rd->part_cap) return -2;
    const int fused = nb <= g_fuse_max_blocks;
    int *cnt = rd->cnt + BBJD_CNT_MID;
#define BBJD_LAUNCH_MID(KW)                                                                                         \
    bbjd::gl<&k_cgs_mid_b<R, KW>, TPB>(k_cgs_mid<R, KW>, dim3((unsigned)nb), 0, s, (i64)n, k,                        \
                                       (const Cx<R> *)Q,                                                             \
                       (i64)ldq, (const Cx<R> *)f.h1, (Cx<R> *)v, (const Cx<R> *)xa, (const Cx<R> *)xb, ca, cb,      \
                       (int)per, (Cx<R> *)rd->part, (R *)rd->pnorm, cnt, f, fused)
    switch (cgs_kw(k)) {
    case 2: BBJD_LAUNCH_MID(2); break;
    case 4: BBJD_LAUNCH_MID(4); break;
    case 8: BBJD_LAUNCH_MID(8); break;
    case 16: BBJD_LAUNCH_MID(16); break;
    case 32: BBJD_LAUNCH_MID(32); break;
    default: return -2;
    }
#undef BBJD_LAUNCH_MID
    if (!fused)
        bbjd::gl<&k_cgs_mid_fin_b<R>, TPB>(k_cgs_mid_fin<R>, dim3((unsigned)((k + 3) / 4)), 0, s, (int)nb, k,
                           (const Cx<R> *)rd->part, (const R *)rd->pnorm, f);
    return chk();
}

// Last launch of the fused CGS2 (+ the norm finish when the grid is large).
template <typename R>
static int cgs_final(hipStream_t s, i64 n, int k, const void *Q, i64 ldq, void *v, bbjd_red *rd, NormFin fin,
                     int scale, int force)
{
    constexpr int RW = 2;
    const int nt = ntiles(n, 64 * RW);
    const int fused = nt <= g_fuse_max_blocks;
    // mode 2 never trusts the estimate (e <= ||v1||^2 < 2 ||v1||^2)
    const double trust = g_cgs_fused == 2 ? 2.0 : 0.25;
    bbjd::gl<&k_cgs_final_b<R, RW>, TPB>(k_cgs_final<R, RW>, dim3(nt), 0, s, (i64)n, k, (const Cx<R> *)Q, (i64)ldq,
                       (const Cx<R> *)rd->h2, (Cx<R> *)v, scale, force, trust, (const double *)rd->scal,
                       rd->flag, (R *)rd->pnorm, rd->cnt + BBJD_CNT_UPD, fin, fused);
    if (!fused && fin.mode >= 0)
        bbjd::gl<&k_norm_fin_b<R>, TPB>(k_norm_fin<R>, dim3(1), 0, s, (const R *)rd->pnorm, nt, fin,
                           force ? (const int *)nullptr : (const int *)(rd->flag + FL_PASS2));
    return chk();
}

// norm finish of the fused CGS2's last pass; with unit output it also does
// the rare rescale by the computed norm
static inline NormFin fin_unit(bbjd_red *rd, int slot, void *x, i64 n, int scale)
{
    NormFin f = fin_norm(rd, scale ? 3 : 0, slot);
    f.x = x;
    f.n = n;
    return f;
}

static inline bool cgs_fusable(int k, const bbjd_red *rd)
{
    return g_cgs_fused && cgs_kw(k) > 0 && (i64)k <= rd->part_cap;
}

// CGS2 with the DGKS test.  k <= 128: the fused form (three reads of Q when
// the second pass runs).  Otherwise four launches: projection (+ ||v||
// before), update (+ ||v|| after, decision), and the second pass, which
// reads the decision flag on the device and returns at once when it is not
// needed.  No host round trip either way.  scale: normalise the result when
// its norm exceeds 1e-300 (scal[out_slot] keeps the norm before scaling).
template <typename R>
static int cgs2(hipStream_t s, i64 n, int k, const void *Q, i64 ldq, void *v, bbjd_red *rd, int out_slot,
                void *hsum, int scale = 0)
{
    int rc = 0;
    if (k == 0) {
        rc |= norm_into<R>(s, n, v, rd, out_slot);
        if (scale) rc |= scale_inv<R>(s, n, v, rd->scal, out_slot, nullptr);
        return rc;
    }
    if (cgs_fusable(k, rd)) {
        rc |= proj_pass<R>(s, n, k, Q, ldq, v, nullptr, rd->h, nullptr, rd, SL_NB, nullptr);
        MidFin f;
        f.scal = rd->scal; f.flag = rd->flag; f.out_slot = out_slot; f.decide = 1;
        f.h1 = rd->h; f.h2 = rd->h2; f.hsum = hsum;
        rc |= cgs_mid<R>(s, n, k, Q, ldq, v, rd, f);
        rc |= cgs_final<R>(s, n, k, Q, ldq, v, rd, fin_unit(rd, out_slot, v, n, scale), scale, 0);
        return rc;
    }
    rc |= proj_pass<R>(s, n, k, Q, ldq, v, nullptr, rd->h, nullptr, rd, SL_NB, nullptr);
    NormFin f1 = fin_norm(rd, 1, out_slot);
    if (hsum) { f1.hsum = hsum; f1.h = rd->h; f1.k = k; }
    rc |= update_pass<R>(s, n, k, Q, ldq, rd->h, (R)-1, 1, v, rd, f1, nullptr);
    const int *cond = rd->flag + FL_PASS2;
    rc |= proj_pass<R>(s, n, k, Q, ldq, v, nullptr, rd->h2, nullptr, rd, -1, cond);
    NormFin f2 = fin_norm(rd, 2, out_slot);
    if (hsum) { f2.hsum = hsum; f2.h = rd->h; f2.h2 = rd->h2; f2.k = k; }
    rc |= update_pass<R>(s, n, k, Q, ldq, rd->h2, (R)-1, 1, v, rd, f2, cond);
    if (scale) rc |= scale_inv<R>(s, n, v, rd->scal, out_slot, nullptr);
    return rc;
}

// The test-space expansion (k <= 128): w = unit CGS2 of ca a + cb b against
// W (a = A v, b = B v; the combination is formed in the first two passes,
// never stored before the result), and in the same first pass over W the
// pencil column W^H a, W^H b, then the new diagonal entries w^H a, w^H b
// after the normalisation: colA[0..k], colB[0..k].  Returns 1 without doing
// anything when the shape takes the unfused path (the caller forms w and
// projects separately).
template <typename R>
static int cgs2_unit_pencil(hipStream_t s, i64 n, int k, const void *W, i64 ldw, void *w, const void *a,
                            const void *b, Cx<R> ca, Cx<R> cb, void *colA, void *colB, bbjd_red *rd, int out_slot)
{
    if (k == 0 || !cgs_fusable(k, rd)) return 1;
    const int ctiles = (k + 7) / 8;
    const i64 chunk = proj_chunk(n, ctiles, k);
    if ((n + chunk - 1) / chunk * 3 * (i64)k > rd->part_cap) return 1;
    int rc = 0;
    rc |= proj_pass_nv<R, 3>(s, n, k, W, ldw, nullptr, a, b, rd->h, colA, colB, rd, SL_NB, nullptr, ca, cb);
    MidFin f;
    f.scal = rd->scal; f.flag = rd->flag; f.out_slot = out_slot; f.decide = 1;
    f.h1 = rd->h; f.h2 = rd->h2; f.hsum = nullptr;
    rc |= cgs_mid<R>(s, n, k, W, ldw, w, rd, f, a, b, ca, cb);
    rc |= cgs_final<R>(s, n, k, W, ldw, w, rd, fin_unit(rd, out_slot, w, n, 1), 1, 0);
    rc |= proj_pass_nv<R, 2>(s, n, 1, w, n, a, b, nullptr, (Cx<R> *)colA + k, (Cx<R> *)colB + k, nullptr, rd, -1,
                             nullptr);
    return rc ? -1 : 0;
}

// v -= Q Q^H v twice (no norm test); fused as in cgs2 for k <= 128.
template <typename R>
static int project_out(hipStream_t s, i64 n, int k, const void *Q, i64 ldq, void *v, bbjd_red *rd)
{
    if (k == 0) return 0;
    int rc = 0;
    if (cgs_fusable(k, rd)) {
        rc |= proj_pass<R>(s, n, k, Q, ldq, v, nullptr, rd->h, nullptr, rd, -1, nullptr);
        MidFin f;
        f.scal = rd->scal; f.flag = rd->flag; f.out_slot = 0; f.decide = 0;
        f.h1 = rd->h; f.h2 = rd->h2; f.hsum = nullptr;
        rc |= cgs_mid<R>(s, n, k, Q, ldq, v, rd, f);
        rc |= cgs_final<R>(s, n, k, Q, ldq, v, rd, fin_none(), 0, 1);
        return rc;
    }
    for (int pass = 0; pass < 2; pass++) {
        rc |= proj_pass<R>(s, n, k, Q, ldq, v, nullptr, rd->h, nullptr, rd, -1, nullptr);
        rc |= update_pass<R>(s, n, k, Q, ldq, rd->h, (R)-1, 1, v, rd, fin_none(), nullptr);
    }
    return rc;
}

template <typename R>
static int gemv_n(hipStream_t s, i64 n, int k, const void *Q, i64 ldq, const void *c, void *y, bbjd_red *rd,
                  int slot)
{
    return update_pass<R>(s, n, k, Q, ldq, c, (R)1, 0, y, rd, slot >= 0 ? fin_norm(rd, 0, slot) : fin_none(),
                          nullptr);
}

template <typename R>
static int lincomb(hipStream_t s, i64 n, double are, double aim, const void *x, double bre, double bim,
                   const void *y, void *out, bbjd_red *rd, int slot, const double *co = nullptr)
{
    const int rows = vec_rows(n);
    const int nt = ntiles(n, rows);
    const int fused = nt <= g_fuse_max_blocks;
#define BBJD_LAUNCH_LC(E)                                                                                          \
    bbjd::gl<&k_lincomb_b<R, E>, TPB>(k_lincomb<R, E>, dim3(nt), 0, s, (i64)n, (R)are, (R)aim, (const Cx<R> *)x,      \
                       (R)bre, (R)bim, (const Cx<R> *)y, (Cx<R> *)out, rd ? (R *)rd->pnorm : (R *)nullptr,          \
                       rd ? rd->cnt + BBJD_CNT_VEC : (int *)nullptr, rd ? fin_norm(rd, 0, slot) : fin_none(), fused, \
                       co)
    BBJD_E_SWITCH(rows, BBJD_LAUNCH_LC)
#undef BBJD_LAUNCH_LC
    if (rd && !fused)
        bbjd::gl<&k_norm_fin_b<R>, TPB>(k_norm_fin<R>, dim3(1), 0, s, (const R *)rd->pnorm, nt,
                           fin_norm(rd, 0, slot), (const int *)nullptr);
    return chk();
}

// Copy of nw 8-byte words (small device-to-device staging copies that a
// grouped launch can merge).
__device__ __forceinline__ void k_copy_words_b(const VB vb, i64 nw, const unsigned long long *__restrict__ src,
                                               unsigned long long *__restrict__ dst)
{
    for (i64 i = (i64)vb.bx * TPB + threadIdx.x; i < nw; i += (i64)vb.gx * TPB) dst[i] = src[i];
}
__global__ __launch_bounds__(TPB) void k_copy_words(i64 nw, const unsigned long long *__restrict__ src,
                                                    unsigned long long *__restrict__ dst)
{
    k_copy_words_b(VB_HW, nw, src, dst);
}

namespace bbjd {
// The order of a flush.  nj lists of launches, launch i of list j with key
// key(j, i) (0: a closure, launched on its own): each step launches, as one
// group, every list whose next launch has the same key as the next launch
// of the list with the most launches left (ties: the key that heads the
// most lists).  Advancing the longest list first lines the lists' ends up,
// so solves whose launch sequences differ in the middle (preconditioner
// levels of different n) merge again in the common part that follows;
// closures run as soon as they head their list.  emit(sel, n, at) gets the
// n lists of a launch and their positions.  Returns 0.
template <class Len, class Key, class Emit>
int plan_merge(int nj, Len len, Key key, Emit emit)
{
    std::vector<size_t> h(nj, 0);
    int sel[kGroupMax];
    size_t at[kGroupMax];
    for (;;) {
        for (int j = 0; j < nj; j++)
            while (h[j] < len(j) && key(j, h[j]) == 0) {
                sel[0] = j;
                at[0] = h[j];
                emit(sel, 1, at);
                h[j]++;
            }
        int lead = -1;
        size_t lead_rem = 0;
        int lead_cnt = 0;
        for (int j = 0; j < nj; j++) {
            if (h[j] >= len(j)) continue;
            const size_t rem = len(j) - h[j];
            const uintptr_t k = key(j, h[j]);
            int cnt = 0;
            for (int i = 0; i < nj; i++) cnt += h[i] < len(i) && key(i, h[i]) == k;
            if (lead < 0 || rem > lead_rem || (rem == lead_rem && cnt > lead_cnt)) {
                lead = j;
                lead_rem = rem;
                lead_cnt = cnt;
            }
        }
        if (lead < 0) return 0;
        const uintptr_t k = key(lead, h[lead]);
        int n = 0;
        sel[n] = lead;
        at[n++] = h[lead];
        for (int j = 0; j < nj && n < kGroupMax; j++)
            if (j != lead && h[j] < len(j) && key(j, h[j]) == k) {
                sel[n] = j;
                at[n++] = h[j];
            }
        emit(sel, n, at);
        for (int g = 0; g < n; g++) h[sel[g]]++;
    }
}

Recorder *&tl_rec()
{
    static thread_local Recorder *r = nullptr;
    return r;
}
}  // namespace bbjd

extern "C" {

void bbjd_set_fuse_blocks(int max_blocks) { g_fuse_max_blocks = max_blocks; }
void bbjd_set_cgs_mid_blocks(int max_blocks) { g_mid_blocks = max_blocks < 1 ? 512 : max_blocks; }

int bbjd_set_tile_table(int nn, const int64_t *n_upper, int nk, const int *k_upper, const int *proj_blocks,
                        const int *upd_rows)
{
    if (nn == 0 || nk == 0) { g_tt_nn = g_tt_nk = 0; return 0; }
    if (nn < 0 || nk < 0 || nn > TT_MAX || nk > TT_MAX || !n_upper || !k_upper || !proj_blocks || !upd_rows)
        return -2;
    for (int i = 0; i < nn * nk; i++) {
        const int r = upd_rows[i];
        if (proj_blocks[i] <= 0 || !(r == 0 || r == 64 || r == 128 || r == TPB || r == 2 * TPB || r == 4 * TPB))
            return -2;
    }
    for (int i = 0; i < nn; i++) g_tt_n[i] = n_upper[i];
    for (int j = 0; j < nk; j++) g_tt_k[j] = k_upper[j];
    for (int i = 0; i < nn; i++)
        for (int j = 0; j < nk; j++) {
            g_tt_proj[i * TT_MAX + j] = proj_blocks[i * nk + j];
            g_tt_upd[i * TT_MAX + j] = upd_rows[i * nk + j];
        }
    g_tt_nn = nn;
    g_tt_nk = nk;
    return 0;
}

int bbjd_set_tiles(int proj_blocks, int upd_rows, int blocks_target)
{
    auto ok = [](int r) { return r == 0 || r == 64 || r == 128 || r == TPB || r == 2 * TPB || r == 4 * TPB; };
    if (proj_blocks < 0 || !ok(upd_rows) || blocks_target < 0) return -2;
    if (proj_blocks > 0) g_proj_blocks = proj_blocks;
    g_upd_rows = upd_rows;
    if (blocks_target > 0) g_blocks_target = blocks_target;
    return 0;
}

void bbjd_ktime_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_kt_mu);
    g_kt_on = on != 0;
}

// Waits for the recorded launches; returns their count and fills the sums of
// durations (ms) and algorithmic bytes, then clears the record.
int bbjd_ktime_harvest(double *total_ms, double *total_bytes)
{
    std::vector<KRec> recs;
    {
        std::lock_guard<std::mutex> lk(g_kt_mu);
        recs.swap(g_kt);
    }
    double ms = 0, by = 0;
    int cnt = 0;
    for (auto &r : recs) {
        float t = 0;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
            ms += t; by += r.bytes; cnt++;
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_kt_mu);
        for (auto &r : recs) { g_kt_pool.push_back(r.a); g_kt_pool.push_back(r.b); }
    }
    if (total_ms) *total_ms = ms;
    if (total_bytes) *total_bytes = by;
    return cnt;
}

int bbjd_k_cgs2(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, void *v, bbjd_red *rd,
                int out_slot, void *hsum)
{ return cgs2<double>(s, n, k, Q, ldq, v, rd, out_slot, hsum); }

int bbjd_k_cgs2_unit(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, void *v, bbjd_red *rd,
                     int out_slot)
{ return cgs2<double>(s, n, k, Q, ldq, v, rd, out_slot, nullptr, 1); }

int bbjd_k_cgs2_unit_pencil(hipStream_t s, int64_t n, int k, const void *W, int64_t ldw, void *w, const void *a,
                            const void *b, double are, double aim, double bre, double bim, void *colA, void *colB,
                            bbjd_red *rd, int out_slot)
{
    return cgs2_unit_pencil<double>(s, n, k, W, ldw, w, a, b, Cx<double>{are, aim}, Cx<double>{bre, bim}, colA,
                                    colB, rd, out_slot);
}

void bbjd_set_cgs_fused(int mode) { g_cgs_fused = mode == 2 ? 2 : mode != 0; }

int bbjd_k_cgs2_c(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, void *v, bbjd_red *rd,
                  int out_slot, void *hsum)
{ return cgs2<float>(s, n, k, Q, ldq, v, rd, out_slot, hsum); }

int bbjd_k_project_out(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, void *v, bbjd_red *rd)
{ return project_out<double>(s, n, k, Q, ldq, v, rd); }

int bbjd_k_project_out_c(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, void *v, bbjd_red *rd)
{ return project_out<float>(s, n, k, Q, ldq, v, rd); }

int bbjd_k_gemv_c(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, const void *x, const void *x2,
                  void *out, void *out2, bbjd_red *rd)
{ return proj_pass<double>(s, n, k, Q, ldq, x, x2, out, out2, rd, -1, nullptr); }

int bbjd_k_gemv_c_c(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, const void *x, const void *x2,
                    void *out, void *out2, bbjd_red *rd)
{ return proj_pass<float>(s, n, k, Q, ldq, x, x2, out, out2, rd, -1, nullptr); }

// x -= Y[:, :k] (M c), complex64 (M k x k column-major, c k entries; device)
int bbjd_k_update_pre_c(hipStream_t s, int64_t n, int k, const void *Y, int64_t ldy, const void *M, const void *c,
                        void *x)
{
    if (k <= 0) return 0;
    if (k > 2048 || n <= 0) return -2;          // LDS: 2 k complex64
    const size_t lds = 2 * (size_t)k * sizeof(Cx<float>);
    // enough blocks to spread the rows over the device: four rows per
    // thread only on the largest vectors
    const int e = n >= (i64)512 * 4 * TPB ? 4 : n >= (i64)256 * 2 * TPB ? 2 : 1;
    const int nb = (int)((n + (i64)TPB * e - 1) / ((i64)TPB * e));
#define BBJD_LAUNCH_PRE(E)                                                                                         \
    BBJD_SL(s, (k_update_pre<float, E>), dim3(nb), dim3(TPB), lds, (i64)n, k, (const Cx<float> *)Y,     \
                       (i64)ldy, (const Cx<float> *)M, (const Cx<float> *)c, (Cx<float> *)x)
    switch (e) {
    case 4: BBJD_LAUNCH_PRE(4); break;
    case 2: BBJD_LAUNCH_PRE(2); break;
    default: BBJD_LAUNCH_PRE(1); break;
    }
#undef BBJD_LAUNCH_PRE
    return chk();
}

int bbjd_k_gemv_n(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, const void *c, void *y,
                  bbjd_red *rd, int slot)
{ return gemv_n<double>(s, n, k, Q, ldq, c, y, rd, slot); }

int bbjd_k_gemv_n_host(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, const void *hc, void *y)
{
    if (k < 1 || k > 128) return -2;
    Coef128 c;
    memcpy(c.v, hc, (size_t)k * sizeof(Cx<double>));
    BBJD_SL(s, k_gemv_n_karg<2>, dim3(ntiles(n, 128)), dim3(TPB), 0, (i64)n, k, (const Cx<double> *)Q,
                       (i64)ldq, c, (Cx<double> *)y);
    return chk();
}

int bbjd_k_gemv_n_c(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, const void *c, void *y,
                    bbjd_red *rd, int slot)
{ return gemv_n<float>(s, n, k, Q, ldq, c, y, rd, slot); }

int bbjd_k_norm(hipStream_t s, int64_t n, const void *x, bbjd_red *rd, int slot)
{ return norm_into<double>(s, n, x, rd, slot); }

int bbjd_k_norm_c(hipStream_t s, int64_t n, const void *x, bbjd_red *rd, int slot)
{ return norm_into<float>(s, n, x, rd, slot); }

int bbjd_k_scale_inv(hipStream_t s, int64_t n, void *x, const double *scal, int slot, int mode)
{
    return bbjd_k_scale_inv2(s, n, x, nullptr, scal, slot, mode);
}

int bbjd_k_scale_inv_c(hipStream_t s, int64_t n, void *x, const double *scal, int slot, int mode)
{
    return bbjd_k_scale_inv2_c(s, n, x, nullptr, scal, slot, mode);
}

int bbjd_k_scale_inv2(hipStream_t s, int64_t n, void *x, void *y, const double *scal, int slot, int mode)
{
    bbjd::gl<&k_scale_inv_b<double>, TPB>(k_scale_inv<double>, dim3(grid_stride_blocks(n)), 0, s, (i64)n,
                       (Cx<double> *)x, scal, slot, mode, (const int *)nullptr, (Cx<double> *)y);
    return chk();
}

int bbjd_k_scale_inv2_c(hipStream_t s, int64_t n, void *x, void *y, const double *scal, int slot, int mode)
{
    bbjd::gl<&k_scale_inv_b<float>, TPB>(k_scale_inv<float>, dim3(grid_stride_blocks(n)), 0, s, (i64)n,
                       (Cx<float> *)x, scal, slot, mode, (const int *)nullptr, (Cx<float> *)y);
    return chk();
}

int bbjd_k_scale(hipStream_t s, int64_t n, void *x, double are, double aim)
{
    bbjd::gl<&k_scale_b<double>, TPB>(k_scale<double>, dim3(grid_stride_blocks(n)), 0, s, (i64)n, (Cx<double> *)x,
                       are, aim);
    return chk();
}

int bbjd_k_scale_c(hipStream_t s, int64_t n, void *x, double are, double aim)
{
    bbjd::gl<&k_scale_b<float>, TPB>(k_scale<float>, dim3(grid_stride_blocks(n)), 0, s, (i64)n, (Cx<float> *)x,
                       (float)are, (float)aim);
    return chk();
}

int bbjd_k_lincomb(hipStream_t s, int64_t n, double are, double aim, const void *x, double bre, double bim,
                   const void *y, void *out)
{ return lincomb<double>(s, n, are, aim, x, bre, bim, y, out, nullptr, 0); }

int bbjd_k_lincomb_c(hipStream_t s, int64_t n, double are, double aim, const void *x, double bre, double bim,
                     const void *y, void *out)
{ return lincomb<float>(s, n, are, aim, x, bre, bim, y, out, nullptr, 0); }

int bbjd_k_lincomb_dev(hipStream_t s, int64_t n, const double *co, const void *x, const void *y, void *out)
{ return lincomb<double>(s, n, 0, 0, x, 0, 0, y, out, nullptr, 0, co); }

int bbjd_k_lincomb_dev_c(hipStream_t s, int64_t n, const double *co, const void *x, const void *y, void *out)
{ return lincomb<float>(s, n, 0, 0, x, 0, 0, y, out, nullptr, 0, co); }

int bbjd_k_lincomb_norm(hipStream_t s, int64_t n, double are, double aim, const void *x, double bre, double bim,
                        const void *y, void *r, bbjd_red *rd, int slot)
{ return lincomb<double>(s, n, are, aim, x, bre, bim, y, r, rd, slot); }

int bbjd_k_svdv(hipStream_t s, int64_t n, const void *u, void *z, const void *sv, bbjd_red *rd)
{
    Cx<double> *d = (Cx<double> *)rd->coef2;
    int left = 0;
    int rc = proj_pass_nv<double, 2>(s, n, 1, u, n, z, sv, nullptr, d, d + 1, nullptr, rd, -1, nullptr,
                                     Cx<double>{0, 0}, Cx<double>{0, 0}, g_svdv_fused ? &left : nullptr);
    if (left)
        BBJD_SL(s, k_axpy_svdv_part, dim3(grid_stride_blocks(n)), dim3(TPB), 0, (i64)n,
                           (const Cx<double> *)rd->part, left, d, (const Cx<double> *)sv, (Cx<double> *)z);
    else
        BBJD_SL(s, k_axpy_svdv, dim3(grid_stride_blocks(n)), dim3(TPB), 0, (i64)n,
                           (const Cx<double> *)d, (const Cx<double> *)sv, (Cx<double> *)z);
    return rc | chk();
}

void bbjd_set_svdv_fused(int on) { g_svdv_fused = on != 0; }

void bbjd_set_gram_blocks(int blocks) { if (blocks > 0) g_gram_blocks = blocks; }

void bbjd_set_rotate_mfma(int on) { g_rotate_mfma = on < 0 ? 0 : (on > 2 ? 2 : on); }
void bbjd_set_rotate_chunk(int kc) { g_rot_kc = (kc == 8 || kc == 16 || kc == 32) ? kc : 0; }

// G (p x q, column-major, ld ldg) = X^H Y; partials in rd->part.
int bbjd_k_gram(hipStream_t s, int64_t n, int p, const void *X, int64_t ldx, int q, const void *Y, int64_t ldy,
                void *G, int64_t ldg, bbjd_red *rd)
{
    if (p <= 0 || q <= 0) return 0;
    if (n <= 0 || ldg < p) return -2;
    const i64 pq = (i64)p * q;
    const int tiles = ((p + GR_T - 1) / GR_T) * ((q + GR_T - 1) / GR_T);
    i64 nch = (g_gram_blocks + tiles - 1) / tiles;
    if (nch * pq > rd->part_cap) nch = rd->part_cap / pq;
    if (nch < 1) return -2;
    // at least 256 rows per chunk: the partials stay small beside the inputs
    const i64 maxch = (n + 255) / 256;
    if (nch > maxch) nch = maxch;
    i64 chunk = (n + nch - 1) / nch;
    chunk = (chunk + GR_SR - 1) / GR_SR * GR_SR;
    nch = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_gram, dim3(tiles, (unsigned)nch), dim3(TPB), 0, s, (i64)n, chunk, p, q,
                       (const Cx<double> *)X, (i64)ldx, (const Cx<double> *)Y, (i64)ldy, (Cx<double> *)rd->part);
    hipLaunchKernelGGL(k_gram_finish, dim3((unsigned)((pq * 16 + TPB - 1) / TPB)), dim3(TPB), 0, s, (int)nch, p, q,
                       (const Cx<double> *)rd->part, (Cx<double> *)G, (i64)ldg);
    return chk();
}

int bbjd_k_rotate(hipStream_t s, int64_t n, int k, int keep, void *X, int64_t ldx, const void *Zr)
{
    if (keep > k || keep <= 0) return -2;
    const int k4 = (k + 3) & ~3;
    // matrix cores, streaming form (any k, keep <= 128): 64-row blocks, the
    // input columns through double-buffered LDS chunks
    if (g_rotate_mfma == 1 && keep <= 128) {
        const dim3 g((unsigned)((n + 63) / 64));
        const int nt = (keep + 15) / 16;
#define BBJD_ROT_S(NT, KC)                                                                                       \
    hipLaunchKernelGGL((k_rotate_mfma_s<NT, KC>), g, dim3(TPB), 0, s, (i64)n, k, keep, (Cx<double> *)X, (i64)ldx, \
                       (const Cx<double> *)Zr)
        const int kc = g_rot_kc;
        if (nt <= 1) BBJD_ROT_S(1, 16);
        else if (nt <= 2) BBJD_ROT_S(2, 16);
        else if (nt <= 4) {
            if (kc == 16) BBJD_ROT_S(4, 16);
            else if (kc == 32) BBJD_ROT_S(4, 32);
            else BBJD_ROT_S(4, 8);
        } else if (nt <= 6) BBJD_ROT_S(6, 8);
        else if (kc == 16) BBJD_ROT_S(8, 16);
        else BBJD_ROT_S(8, 8);
#undef BBJD_ROT_S
        return chk();
    }
    // matrix cores, staged form: 32 rows per block while the staged rows fit
    // 64 KB of LDS and the output has at most 16 blocks of 16 x 16, else 16
    if (g_rotate_mfma && (size_t)32 * k4 * 16 <= 64 * 1024 && 2 * ((keep + 15) / 16) <= 16) {
        hipLaunchKernelGGL((k_rotate_mfma<32>), dim3((unsigned)((n + 31) / 32)), dim3(TPB), (size_t)32 * k4 * 16, s,
                           (i64)n, k, keep, (Cx<double> *)X, (i64)ldx, (const Cx<double> *)Zr);
        return chk();
    }
    if (g_rotate_mfma && (size_t)16 * k4 * 16 <= 64 * 1024 && (keep + 15) / 16 <= 16) {
        hipLaunchKernelGGL((k_rotate_mfma<16>), dim3((unsigned)((n + 15) / 16)), dim3(TPB), (size_t)16 * k4 * 16, s,
                           (i64)n, k, keep, (Cx<double> *)X, (i64)ldx, (const Cx<double> *)Zr);
        return chk();
    }
    int rt = 16;
    while (rt > 4 && (size_t)rt * (k + keep) * 16 > 64 * 1024) rt /= 2;
    if ((size_t)rt * (k + keep) * 16 > 64 * 1024) return -2;
    const size_t lds = (size_t)rt * (k + keep) * 16;
    const int nb = (int)((n + rt - 1) / rt);
    if (rt == 16)
        hipLaunchKernelGGL((k_rotate<16>), dim3(nb), dim3(TPB), lds, s, (i64)n, k, keep, (Cx<double> *)X, (i64)ldx,
                           (const Cx<double> *)Zr);
    else if (rt == 8)
        hipLaunchKernelGGL((k_rotate<8>), dim3(nb), dim3(TPB), lds, s, (i64)n, k, keep, (Cx<double> *)X, (i64)ldx,
                           (const Cx<double> *)Zr);
    else
        hipLaunchKernelGGL((k_rotate<4>), dim3(nb), dim3(TPB), lds, s, (i64)n, k, keep, (Cx<double> *)X, (i64)ldx,
                           (const Cx<double> *)Zr);
    return chk();
}

int bbjd_k_z2c(hipStream_t s, int64_t n, const void *z, void *c)
{
    bbjd::gl<&k_z2c_b, TPB>(k_z2c, dim3(grid_stride_blocks(n)), 0, s, (i64)n, (const Cx<double> *)z,
                       (Cx<float> *)c);
    return chk();
}

int bbjd_k_c2z(hipStream_t s, int64_t n, const void *c, void *z)
{
    bbjd::gl<&k_c2z_b, TPB>(k_c2z, dim3(grid_stride_blocks(n)), 0, s, (i64)n, (const Cx<float> *)c,
                       (Cx<double> *)z);
    return chk();
}

// ---------------------------------------------------------------------------
// Recording and grouped launch (launch.h).  bbjd_rec_flush launches the jobs'
// lists in the order bbjd::plan_merge gives.  Each job's launches keep their
// order (one stream), and jobs share no writable memory, so the interleaving
// does not change any job's results.
void *bbjd_rec_create(void) { return new bbjd::Recorder; }
void bbjd_rec_destroy(void *rec) { delete static_cast<bbjd::Recorder *>(rec); }
void bbjd_rec_begin(void *rec, int job)
{
    bbjd::Recorder *r = static_cast<bbjd::Recorder *>(rec);
    if ((int)r->jobs.size() <= job) r->jobs.resize(job + 1);
    r->cur = job;
    bbjd::tl_rec() = r;
}
void bbjd_rec_end(void) { bbjd::tl_rec() = nullptr; }
void bbjd_rec_counts(const void *rec, long *recorded, long *launches)
{
    const bbjd::Recorder *r = static_cast<const bbjd::Recorder *>(rec);
    *recorded = r->recorded.load(std::memory_order_relaxed);
    *launches = r->launches.load(std::memory_order_relaxed);
}

int bbjd_rec_flush(void *rec, hipStream_t s, long *launches)
{
    bbjd::Recorder *r = static_cast<bbjd::Recorder *>(rec);
    const int nj = (int)r->jobs.size();
    long nl = 0;
    const int rc = bbjd::plan_merge(
        nj, [&](int j) { return r->jobs[j].size(); },
        [&](int j, size_t i) { return (uintptr_t)r->jobs[j][i].fn; },
        [&](const int *sel, int n, const size_t *at) {
            const bbjd::Op &o0 = r->jobs[sel[0]][at[0]];
            if (!o0.fn) {
                o0.solo(s);
            } else {
                bbjd::Op *ops[bbjd::kGroupMax];
                for (int g = 0; g < n; g++) ops[g] = &r->jobs[sel[g]][at[g]];
                o0.fn(ops, n, s, r->arena.data());
            }
            nl++;
        });
    for (auto &l : r->jobs) l.clear();
    r->arena.clear();
    r->launches.fetch_add(nl, std::memory_order_relaxed);
    if (launches) *launches = nl;
    if (rc || hipGetLastError() != hipSuccess) return -1;
    return 0;
}

int bbjd_rec_plan(int nj, const int *lens, const int64_t *keys, int *groups, int cap)
{
    if (nj < 0 || (nj && (!lens || !keys || !groups))) return -1;
    std::vector<size_t> off(nj + 1, 0);
    for (int j = 0; j < nj; j++) off[j + 1] = off[j] + (size_t)lens[j];
    int used = 0, nl = 0, over = 0;
    const int rc = bbjd::plan_merge(
        nj, [&](int j) { return (size_t)lens[j]; }, [&](int j, size_t i) { return (uintptr_t)keys[off[j] + i]; },
        [&](const int *sel, int n, const size_t *) {
            if (used + 1 + n > cap) { over = 1; return; }
            groups[used++] = n;
            for (int g = 0; g < n; g++) groups[used++] = sel[g];
            nl++;
        });
    return rc || over ? -1 : nl;
}

int bbjd_copy_async(void *dst, const void *src, size_t bytes, int kind, hipStream_t s)
{
    if (!bbjd::tl_rec()) return hipMemcpyAsync(dst, src, bytes, (hipMemcpyKind)kind, s) == hipSuccess ? 0 : -1;
    bbjd::sl(s, [=](hipStream_t s_) { (void)hipMemcpyAsync(dst, src, bytes, (hipMemcpyKind)kind, s_); });
    return 0;
}

int bbjd_memset_async(void *dst, int value, size_t bytes, hipStream_t s)
{
    if (!bbjd::tl_rec()) return hipMemsetAsync(dst, value, bytes, s) == hipSuccess ? 0 : -1;
    bbjd::sl(s, [=](hipStream_t s_) { (void)hipMemsetAsync(dst, value, bytes, s_); });
    return 0;
}

int bbjd_k_copy(hipStream_t s, void *dst, const void *src, size_t bytes)
{
    if (bytes % 8) return -2;
    const i64 nw = (i64)(bytes / 8);
    bbjd::gl<&k_copy_words_b, TPB>(k_copy_words, dim3((unsigned)((nw + TPB - 1) / TPB)), 0, s, nw,
                                   (const unsigned long long *)src, (unsigned long long *)dst);
    return chk();
}

// ---------------------------------------------------------------------------
// Kernel benchmark: times one kernel in isolation with HIP events on a private
// stream.  which: 0 k_proj one vector, 1 k_proj two vectors, 2 k_update,
// 3 k_rotate (k -> k/2 columns), 4 full CGS2 (one pass + skipped pass 2),
// 5 a plain copy of n complex values, 6 the two-pass projection (both
// passes always run: the fused form's three reads of Q for k <= 128), 7 the
// Gram matrix Q^H Q (k x k, FP64 matrix cores; 8 n k^2 flops).
// Returns the average ms per launch and the algorithmic bytes per launch.
__global__ void k_fill(i64 cnt, Cx<double> *a, double scale)
{
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < cnt; i += (i64)gridDim.x * TPB)
        a[i] = cx(scale * (double)((i * 2654435761LL) % 1000 - 500) * 1e-3,
                  scale * (double)((i * 40503LL) % 997 - 498) * 1e-3);
}

// Plain streaming copy: the reference point for counter-based HBM traffic.
__global__ void k_copy(i64 cnt, const Cx<double> *__restrict__ x, Cx<double> *__restrict__ y)
{
    for (i64 i = (i64)blockIdx.x * TPB + threadIdx.x; i < cnt; i += (i64)gridDim.x * TPB) y[i] = x[i];
}

int bbjd_kbench(int which, int64_t n, int k, int reps, double *avg_ms, double *bytes)
{
    if (n <= 0 || k <= 0 || k > BBJD_MAXK || reps <= 0) return -2;
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return -1;
    const i64 ld = (n + 63) / 64 * 64;
    void *Q = nullptr, *x = nullptr, *x2 = nullptr, *blk = nullptr, *Zr = nullptr;
    const i64 pe = bbjd_red_part_entries(n, k + 2);
    int rc = 0;
    rc |= hipMalloc(&Q, (size_t)ld * k * 16) != hipSuccess;
    rc |= hipMalloc(&x, (size_t)ld * 16) != hipSuccess;
    rc |= hipMalloc(&x2, (size_t)ld * 16) != hipSuccess;
    const size_t npn = (size_t)ntiles(n, BBJD_TILE_MIN_ROWS) + 64;
    rc |= hipMalloc(&blk, (size_t)pe * 16 + npn * 8 + (size_t)(k + 8) * 16 * 5 + 8192) != hipSuccess;
    rc |= hipMalloc(&Zr, (size_t)k * k * 16) != hipSuccess;
    auto release = [&]() {
        (void)hipFree(Q); (void)hipFree(x); (void)hipFree(x2); (void)hipFree(blk); (void)hipFree(Zr);
    };
    if (rc) { release(); (void)hipStreamDestroy(s); return -1; }
    bbjd_red rd;
    char *b = (char *)blk;
    rd.part = b; b += (size_t)pe * 16;
    rd.pnorm = (double *)b; b += npn * 8;
    rd.h = b; b += (size_t)(k + 8) * 16;
    rd.h2 = b; b += (size_t)(k + 8) * 16;
    rd.hsum = b; b += (size_t)(k + 8) * 16;
    rd.coef = b; b += (size_t)(k + 8) * 16;
    rd.coef2 = b; b += (size_t)(k + 8) * 16;
    rd.scal = (double *)b; rd.flag = (int *)(b + 256); rd.cnt = (int *)(b + 512);
    rd.part_cap = pe;
    hipLaunchKernelGGL(k_fill, dim3(4096), dim3(TPB), 0, s, (i64)ld * k, (Cx<double> *)Q, 1.0);
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(TPB), 0, s, (i64)ld, (Cx<double> *)x, 0.7);
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(TPB), 0, s, (i64)ld, (Cx<double> *)x2, 0.3);
    hipLaunchKernelGGL(k_fill, dim3(64), dim3(TPB), 0, s, (i64)k * k, (Cx<double> *)Zr, 0.1);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(TPB), 0, s, (i64)(k + 8), (Cx<double> *)rd.h, 0.01);
    (void)hipMemsetAsync(rd.flag, 0, 256 + BBJD_CNT_SLOTS * 4, s);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    double by = 0;
    auto launch = [&]() {
        switch (which) {
        case 0:
            (void)proj_pass<double>(s, n, k, Q, ld, x, nullptr, rd.h2, nullptr, &rd, -1, nullptr);
            by = 16.0 * n * (k + 1.0);
            break;
        case 1:
            (void)proj_pass<double>(s, n, k, Q, ld, x, x2, rd.h2, rd.hsum, &rd, -1, nullptr);
            by = 16.0 * n * (k + 2.0);
            break;
        case 2:
            (void)update_pass<double>(s, n, k, Q, ld, rd.h, -1.0, 1, x, &rd, fin_norm(&rd, 0, SL_T1), nullptr);
            by = 16.0 * n * (k + 2);
            break;
        case 5:
            hipLaunchKernelGGL(k_copy, dim3(2048), dim3(TPB), 0, s, (i64)n, (const Cx<double> *)x,
                               (Cx<double> *)x2);
            by = 32.0 * n;
            break;
        case 6:
            (void)project_out<double>(s, n, k, Q, ld, x, &rd);
            // fused (k <= 128): three reads of Q, x read three times and
            // written twice; otherwise four passes of k + 2
            by = cgs_fusable(k, &rd) ? 16.0 * n * (3.0 * k + 5) : 16.0 * n * (4.0 * k + 6);
            break;
        case 7:
            // G = Q^H Q on the matrix cores; flops 8 n k^2 (reported by the
            // caller from the shape), bytes: Q staged twice
            (void)bbjd_k_gram(s, n, k, Q, ld, k, Q, ld, Zr, k, &rd);
            by = 16.0 * n * 2.0 * k;
            break;
        case 3:
            bbjd_k_rotate(s, n, k, k / 2 > 0 ? k / 2 : 1, Q, ld, Zr);
            by = 16.0 * n * (k + (k / 2 > 0 ? k / 2 : 1));
            break;
        default:
            cgs2<double>(s, n, k, Q, ld, x, &rd, SL_T0, nullptr);
            by = 16.0 * n * (2.0 * k + 3);
            break;
        }
    };
    launch();
    launch();
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; r++) launch();
    (void)hipEventRecord(e1, s);
    float ms = 0;
    rc |= hipEventSynchronize(e1) != hipSuccess;
    rc |= hipEventElapsedTime(&ms, e0, e1) != hipSuccess;
    rc |= hipGetLastError() != hipSuccess;
    if (avg_ms) *avg_ms = ms / reps;
    if (bytes) *bytes = by;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    release();
    (void)hipStreamDestroy(s);
    return rc ? -1 : 0;
}

// The inner solve's projection Q^H x (complex64, k columns: k_proj plus
// its finish) of `jobs` independent solves, each with its own Q, x and
// reduction scratch, recorded and flushed as grouped launches (jobs > 1) or
// launched one after another (jobs = 1 with the same harness), reps times.
// *avg_ms: wall time per rep on the stream (host recording included);
// *bytes: algorithmic bytes per rep, jobs x 8 n (k + 1).  The kernel trace of
// a run gives the grouped kernel's own duration.
int bbjd_kbench_grouped(int jobs, int64_t n, int k, int reps, double *avg_ms, double *bytes)
{
    if (jobs < 1 || jobs > bbjd::kGroupMax || n <= 0 || k <= 0 || k > 128 || reps <= 0) return -2;
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return -1;
    const i64 ld = (n + 63) / 64 * 64;
    const i64 pe = bbjd_red_part_entries(n, k + 2);
    const size_t npn = (size_t)ntiles(n, BBJD_TILE_MIN_ROWS) + 64;
    const size_t per_q = (size_t)ld * k * 8, per_x = (size_t)ld * 8;
    const size_t per_blk = (size_t)pe * 8 + npn * 8 + (size_t)(k + 8) * 8 * 5 + 8192;
    void *mem = nullptr;
    if (hipMalloc(&mem, (per_q + per_x + per_blk) * jobs) != hipSuccess) { (void)hipStreamDestroy(s); return -1; }
    std::vector<bbjd_red> rd(jobs);
    std::vector<char *> Q(jobs), x(jobs);
    char *b = (char *)mem;
    for (int j = 0; j < jobs; j++) {
        Q[j] = b; b += per_q;
        x[j] = b; b += per_x;
        bbjd_red &r = rd[j];
        r.part = b; b += (size_t)pe * 8;
        r.pnorm = (double *)b; b += npn * 8;
        r.h = b; b += (size_t)(k + 8) * 8;
        r.h2 = b; b += (size_t)(k + 8) * 8;
        r.hsum = b; b += (size_t)(k + 8) * 8;
        r.coef = b; b += (size_t)(k + 8) * 8;
        r.coef2 = b; b += (size_t)(k + 8) * 8;
        r.scal = (double *)b; r.flag = (int *)(b + 256); r.cnt = (int *)(b + 512);
        b += 8192;
        r.part_cap = pe;
        hipLaunchKernelGGL(k_fill, dim3(1024), dim3(TPB), 0, s, (i64)ld * k / 2, (Cx<double> *)Q[j], 0.01 * (j + 1));
        hipLaunchKernelGGL(k_fill, dim3(256), dim3(TPB), 0, s, (i64)ld / 2, (Cx<double> *)x[j], 0.7);
        (void)hipMemsetAsync(r.flag, 0, 256 + BBJD_CNT_SLOTS * 4, s);
    }
    void *rec = jobs > 1 ? bbjd_rec_create() : nullptr;
    auto launch = [&]() {
        for (int j = 0; j < jobs; j++) {
            if (rec) bbjd_rec_begin(rec, j);
            (void)proj_pass<float>(s, n, k, Q[j], ld, x[j], nullptr, rd[j].h2, nullptr, &rd[j], -1, nullptr);
            if (rec) bbjd_rec_end();
        }
        if (rec) (void)bbjd_rec_flush(rec, s, nullptr);
    };
    launch();
    launch();
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; r++) launch();
    (void)hipEventRecord(e1, s);
    float ms = 0;
    int rc = hipEventSynchronize(e1) != hipSuccess;
    rc |= hipEventElapsedTime(&ms, e0, e1) != hipSuccess;
    rc |= hipGetLastError() != hipSuccess;
    if (avg_ms) *avg_ms = ms / reps;
    if (bytes) *bytes = 8.0 * (double)n * (k + 1.0) * jobs;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rec) bbjd_rec_destroy(rec);
    (void)hipFree(mem);
    (void)hipStreamDestroy(s);
    return rc ? -1 : 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free CGS2 on device arrays (bbjd_block_orthogonalize_noalloc):
// the two-stage reductions' scratch (partial sums, per-block norms, the
// coefficient vectors, the scalar slots, the decision flags and the
// last-block-done counters) lives in one file-scope device array instead of
// a per-call allocation.  Sized for n <= kNaMaxN and k <= kNaMaxK:
//   partials   kNaPart complex: the projection keeps at most one partial per
//              row chunk and column (chunks ~ blocks per launch / column
//              tiles, so <= 16 x blocks per launch, and the launcher refuses
//              a grid whose partials exceed part_cap), the fused CGS2's
//              middle pass at most 512 blocks x k (k <= 128);
//   norms      one double per row tile of >= 64 rows;
//   h, h2, hsum, coef: k + 8 complex each.
namespace {
constexpr int64_t kNaMaxN = (int64_t)1 << 23;
constexpr int kNaMaxK = 1024;
constexpr int64_t kNaPart = 65536 + 64;
constexpr int64_t kNaPnorm = kNaMaxN / BBJD_TILE_MIN_ROWS + 64;
constexpr int64_t kNaVec = kNaMaxK + 8;
constexpr size_t na_al(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNaOffPart = 0;
constexpr size_t kNaOffPnorm = kNaOffPart + na_al(kNaPart * 16);
constexpr size_t kNaOffH = kNaOffPnorm + na_al(kNaPnorm * 8);
constexpr size_t kNaOffH2 = kNaOffH + na_al(kNaVec * 16);
constexpr size_t kNaOffHsum = kNaOffH2 + na_al(kNaVec * 16);
constexpr size_t kNaOffCoef = kNaOffHsum + na_al(kNaVec * 16);
constexpr size_t kNaOffCoef2 = kNaOffCoef + na_al(kNaVec * 16);
constexpr size_t kNaOffScal = kNaOffCoef2 + na_al(64 * 16);
constexpr size_t kNaOffInts = kNaOffScal + na_al(SL_COUNT * 8);          // flags, then counters
constexpr int kNaInts = FL_COUNT + BBJD_CNT_SLOTS;
constexpr size_t kNaBytes = kNaOffInts + na_al((size_t)kNaInts * 4);

// Per-device address of `scratch`, from hipGetSymbolAddress on the first
// call on each device (the array's name is not a device address in host
// code).
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void *g_na_ptr[kNaMaxDevices];
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// zeroes the flags and counters (they must start at zero; a previous call
// may have been cut short)
__global__ void k_na_clear(int *ints, int cnt)
{
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) ints[i] = 0;
}

__global__ void k_na_copy_norm(const double *src, double *dst) { if (threadIdx.x == 0) *dst = *src; }

static void *na_scratch_ptr()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_na_ptr[dev] = p;
    }
    return g_na_ptr[dev];
}

extern "C" int bbjd_block_orthogonalize_noalloc(int64_t n, int64_t k, const void *Q, int64_t ldq, void *v,
                                                void *h_work, double *norm_out, void *stream)
{
    if (n <= 0 || n > kNaMaxN || k < 0 || k > kNaMaxK || !v || (k > 0 && (!Q || ldq < n))) return -1;
    char *base = (char *)na_scratch_ptr();
    if (!base) return -2;
    hipStream_t s = (hipStream_t)stream;
    bbjd_red rd;
    rd.part = base + kNaOffPart;
    rd.pnorm = (double *)(base + kNaOffPnorm);
    rd.h = base + kNaOffH;
    rd.h2 = base + kNaOffH2;
    rd.hsum = base + kNaOffHsum;
    rd.coef = base + kNaOffCoef;
    rd.coef2 = base + kNaOffCoef2;
    rd.scal = (double *)(base + kNaOffScal);
    rd.flag = (int *)(base + kNaOffInts);
    rd.cnt = rd.flag + FL_COUNT;
    rd.part_cap = kNaPart;
    hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(TPB), 0, s, rd.flag, kNaInts);
    int rc = chk();
    if (!rc) rc = cgs2<double>(s, n, (int)k, Q, ldq, v, &rd, SL_T0, k > 0 ? h_work : nullptr);
    if (!rc && norm_out) {
        hipLaunchKernelGGL(k_na_copy_norm, dim3(1), dim3(64), 0, s, (const double *)rd.scal + SL_T0, norm_out);
        rc = chk();
    }
    return rc ? -2 : 0;
}
