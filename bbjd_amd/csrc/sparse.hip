// sparse.hip -- device operators for general sparse pencils
// (include/bbjd_sparse.h).
//
// A x and B x are CSR matvecs: a group of G lanes (a power of two up to the
// 64-lane wavefront, from the average row length) per row, partial sums
// reduced across the group by shuffles in a fixed order.
//
// With dense rows pinned (bbjd_sparse_border) the factors are those of
// M11 = M[R', R'] and the pinned rows enter through the Schur complement:
// one block per pinned row sums M21 z in a fixed tree, and one pass forms
// y_R = Sinv (x_R - M21 z) in every block and y_R' = z - Y y_R.
//
// M^-1 b = Pc U^-1 L^-1 Pr b from host-computed LU factors: a gather applies
// Pr into a scratch vector, the two triangular solves run in place on it,
// a gather applies Pc into y.  A triangular solve is ordered by level sets:
// row i of L needs the rows j < i of its nonzeros, so level(i) = 1 + the
// largest level among them and the rows of one level are independent.  A
// level with many rows gets a grid of its own; runs of narrow levels go to
// one workgroup that walks them with a barrier between levels, so no block
// ever waits on another one.  Each row's arithmetic: s[i] = (s[i] -
// sum_j F[i, j] s[j]) * (1 / F[i, i]) with the off-diagonal entries summed
// by the lane group as in the matvec.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/bbjd_sparse.h"
#include "engine.h"

typedef long long i64;

namespace {

struct Z { double re, im; };

__device__ __forceinline__ Z zfma(Z acc, Z a, Z b)
{
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
    return acc;
}

template <int G>
__device__ __forceinline__ Z group_sum(Z a)
{
    for (int off = G / 2; off > 0; off >>= 1) {
        a.re += __shfl_down(a.re, off, G);
        a.im += __shfl_down(a.im, off, G);
    }
    return a;
}

constexpr int kTpb = 256;       // matvec and wide-level blocks
constexpr int kNarrowTpb = 1024; // the one-workgroup level walker

template <int G>
__global__ __launch_bounds__(kTpb) void k_csr_mv(i64 n, const i64 *__restrict__ ptr, const int *__restrict__ col,
                                                 const Z *__restrict__ val, const Z *__restrict__ x, Z *__restrict__ y)
{
    const int lane = threadIdx.x % G;
    const i64 groups = (i64)gridDim.x * (kTpb / G);
    for (i64 row = (i64)blockIdx.x * (kTpb / G) + threadIdx.x / G; row < n; row += groups) {
        Z acc{0.0, 0.0};
        const i64 e = ptr[row + 1];
        for (i64 p = ptr[row] + lane; p < e; p += G) acc = zfma(acc, val[p], x[col[p]]);
        acc = group_sum<G>(acc);
        if (lane == 0) y[row] = acc;
    }
}

template <int G>
__device__ __forceinline__ void tri_row(i64 i, int lane, const i64 *__restrict__ ptr, const int *__restrict__ col,
                                        const Z *__restrict__ val, const Z *__restrict__ dinv, Z *s)
{
    Z acc{0.0, 0.0};
    const i64 e = ptr[i + 1];
    for (i64 p = ptr[i] + lane; p < e; p += G) acc = zfma(acc, val[p], s[col[p]]);
    acc = group_sum<G>(acc);
    if (lane == 0) {
        const Z r{s[i].re - acc.re, s[i].im - acc.im};
        const Z d = dinv[i];
        s[i] = Z{r.re * d.re - r.im * d.im, r.re * d.im + r.im * d.re};
    }
}

// one level, rows order[r0 .. r1), over a grid
template <int G>
__global__ __launch_bounds__(kTpb) void k_tri_wide(i64 r0, i64 r1, const int *__restrict__ order,
                                                   const i64 *__restrict__ ptr, const int *__restrict__ col,
                                                   const Z *__restrict__ val, const Z *__restrict__ dinv, Z *s)
{
    const int lane = threadIdx.x % G;
    const i64 groups = (i64)gridDim.x * (kTpb / G);
    for (i64 r = r0 + (i64)blockIdx.x * (kTpb / G) + threadIdx.x / G; r < r1; r += groups)
        tri_row<G>(order[r], lane, ptr, col, val, dinv, s);
}

// levels l0 .. l1 - 1 (rows order[lvl[l] .. lvl[l + 1])) by one workgroup,
// a barrier after each level
template <int G>
__global__ __launch_bounds__(kNarrowTpb) void k_tri_narrow(int l0, int l1, const i64 *__restrict__ lvl,
                                                           const int *__restrict__ order,
                                                           const i64 *__restrict__ ptr, const int *__restrict__ col,
                                                           const Z *__restrict__ val, const Z *__restrict__ dinv, Z *s)
{
    const int lane = threadIdx.x % G;
    for (int l = l0; l < l1; l++) {
        const i64 e = lvl[l + 1];
        for (i64 r = lvl[l] + threadIdx.x / G; r < e; r += kNarrowTpb / G)
            tri_row<G>(order[r], lane, ptr, col, val, dinv, s);
        __syncthreads();
    }
}

__global__ void k_gather(i64 n, const int *__restrict__ idx, const Z *__restrict__ src, Z *__restrict__ dst)
{
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

__global__ void k_copy(i64 n, const Z *__restrict__ src, Z *__restrict__ dst)
{
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Pinned dense rows: t[r] = sum_j M21[r, j] z[j], one block per row, the
// lanes' partial sums reduced in a fixed tree.
__global__ __launch_bounds__(kTpb) void k_border_dots(i64 m, const Z *__restrict__ M21, const Z *__restrict__ z,
                                                      Z *__restrict__ t)
{
    __shared__ double sr[kTpb], si[kTpb];
    const Z *row = M21 + (i64)blockIdx.x * m;
    Z acc = {0.0, 0.0};
    for (i64 j = threadIdx.x; j < m; j += kTpb) acc = zfma(acc, row[j], z[j]);
    sr[threadIdx.x] = acc.re;
    si[threadIdx.x] = acc.im;
    __syncthreads();
    for (int h = kTpb / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sr[threadIdx.x] += sr[threadIdx.x + h];
            si[threadIdx.x] += si[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) t[blockIdx.x] = Z{sr[0], si[0]};
}

// y_R = Sinv (x_R - t), computed by every block (d <= 64), then
// y[rest[i]] = z[i] - sum_j Y[i + j m] y_R[j] and, from block 0, y[R_j] = y_R[j].
__global__ __launch_bounds__(kTpb) void k_border_combine(i64 m, int d, const int *__restrict__ rows,
                                                         const int *__restrict__ rest, const Z *__restrict__ Sinv,
                                                         const Z *__restrict__ Y, const Z *__restrict__ t,
                                                         const Z *__restrict__ xr, const Z *__restrict__ z,
                                                         Z *__restrict__ y)
{
    __shared__ Z rhs[64], y2[64];
    const int tid = threadIdx.x;
    if (tid < d) rhs[tid] = Z{xr[tid].re - t[tid].re, xr[tid].im - t[tid].im};
    __syncthreads();
    if (tid < d) {
        Z a = {0.0, 0.0};
        for (int c = 0; c < d; c++) a = zfma(a, Sinv[tid + (i64)c * d], rhs[c]);
        y2[tid] = a;
        if (blockIdx.x == 0) y[rows[tid]] = a;
    }
    __syncthreads();
    for (i64 i = (i64)blockIdx.x * kTpb + tid; i < m; i += (i64)gridDim.x * kTpb) {
        Z a = {0.0, 0.0};
        for (int j = 0; j < d; j++) a = zfma(a, Y[i + (i64)j * m], y2[j]);
        const Z zi = z[i];
        y[rest[i]] = Z{zi.re - a.re, zi.im - a.im};
    }
}

inline int grid_for(i64 work, int per_block)
{
    i64 b = (work + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (int)b;
}

// lanes per row from the average row length: a power of two in [2, 64]
inline int lanes_for(i64 nnz, i64 rows)
{
    const double avg = rows > 0 ? (double)nnz / (double)rows : 0.0;
    int g = 2;
    while (g < 64 && g < avg) g *= 2;
    return g;
}

// device CSR (row offsets, columns, values)
struct DCsr {
    i64 *ptr = nullptr;
    int *col = nullptr;
    Z *val = nullptr;
    i64 nnz = 0;
    int lanes = 2;
};

// one triangular factor: off-diagonal part, inverse diagonal, level order
// and the launches of its solve
struct Seg { int wide; i64 l0, l1; };
struct DTri {
    DCsr off;
    Z *dinv = nullptr;
    int *order = nullptr;
    i64 *lvl = nullptr;
    i64 nlev = 0;
    std::vector<i64> hlvl;
    std::vector<Seg> segs;
};

}  // namespace

struct bbjd_sparse_op {
    i64 n = 0;
    hipStream_t stream = nullptr;
    void *block = nullptr;
    size_t bytes = 0;
    DCsr A, B;
    int has_B = 0, has_lu = 0;
    DTri L, U;
    int *iperm_r = nullptr, *perm_c = nullptr;
    Z *s = nullptr;
    // pinned dense rows (bbjd_sparse_border): the LU is M11's, of size m
    int has_border = 0, d = 0;
    i64 m = 0;
    int *rows = nullptr, *rest = nullptr;
    Z *M21 = nullptr, *Y = nullptr, *Sinv = nullptr, *t = nullptr, *xr = nullptr, *z = nullptr;
};

static hipStream_t sp_stream(const bbjd_sparse_op *op)
{
    hipStream_t s = (hipStream_t)bbjd_callback_stream();
    return s ? s : op->stream;
}

static int launch_mv(hipStream_t st, i64 n, const DCsr &M, const void *x, void *y)
{
    const int g = grid_for(n * M.lanes, kTpb);
#define BBJD_SP_MV(G)                                                                                        \
    hipLaunchKernelGGL(k_csr_mv<G>, dim3(g), dim3(kTpb), 0, st, n, (const i64 *)M.ptr, (const int *)M.col,   \
                       (const Z *)M.val, (const Z *)x, (Z *)y)
    switch (M.lanes) {
    case 2: BBJD_SP_MV(2); break;
    case 4: BBJD_SP_MV(4); break;
    case 8: BBJD_SP_MV(8); break;
    case 16: BBJD_SP_MV(16); break;
    case 32: BBJD_SP_MV(32); break;
    default: BBJD_SP_MV(64); break;
    }
#undef BBJD_SP_MV
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static int launch_tri(hipStream_t st, const DTri &T, Z *s)
{
    for (const Seg &g : T.segs) {
        if (g.wide) {
            const i64 r0 = T.hlvl[g.l0], r1 = T.hlvl[g.l0 + 1];
            const int grid = grid_for((r1 - r0) * T.off.lanes, kTpb);
#define BBJD_SP_WIDE(G)                                                                                       \
    hipLaunchKernelGGL(k_tri_wide<G>, dim3(grid), dim3(kTpb), 0, st, r0, r1, (const int *)T.order,             \
                       (const i64 *)T.off.ptr, (const int *)T.off.col, (const Z *)T.off.val, (const Z *)T.dinv, s)
            switch (T.off.lanes) {
            case 2: BBJD_SP_WIDE(2); break;
            case 4: BBJD_SP_WIDE(4); break;
            case 8: BBJD_SP_WIDE(8); break;
            case 16: BBJD_SP_WIDE(16); break;
            case 32: BBJD_SP_WIDE(32); break;
            default: BBJD_SP_WIDE(64); break;
            }
#undef BBJD_SP_WIDE
        } else {
#define BBJD_SP_NARROW(G)                                                                                     \
    hipLaunchKernelGGL(k_tri_narrow<G>, dim3(1), dim3(kNarrowTpb), 0, st, (int)g.l0, (int)g.l1,                \
                       (const i64 *)T.lvl, (const int *)T.order, (const i64 *)T.off.ptr, (const int *)T.off.col, \
                       (const Z *)T.off.val, (const Z *)T.dinv, s)
            switch (T.off.lanes) {
            case 2: BBJD_SP_NARROW(2); break;
            case 4: BBJD_SP_NARROW(4); break;
            case 8: BBJD_SP_NARROW(8); break;
            case 16: BBJD_SP_NARROW(16); break;
            case 32: BBJD_SP_NARROW(32); break;
            default: BBJD_SP_NARROW(64); break;
            }
#undef BBJD_SP_NARROW
        }
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

extern "C" {

int64_t bbjd_sparse_levels(bbjd_int n, const int64_t *ptr, const int32_t *col, int lower, int32_t *level)
{
    if (n < 0 || !ptr || !level || (n > 0 && !col)) return -1;
    int64_t nlev = 0;
    for (i64 t = 0; t < n; t++) {
        const i64 i = lower ? t : n - 1 - t;
        int32_t lv = 0;
        for (i64 p = ptr[i]; p < ptr[i + 1]; p++) {
            const i64 j = col[p];
            if (j == i) continue;
            if (lower ? (j > i || j < 0) : (j < i || j >= n)) return -1;
            if (level[j] + 1 > lv) lv = level[j] + 1;
        }
        level[i] = lv;
        if (lv + 1 > nlev) nlev = lv + 1;
    }
    return nlev;
}

}  // extern "C"

// structural checks of a host CSR matrix
static bool csr_ok(const bbjd_csr *M, i64 n, i64 ncols)
{
    if (!M || M->n != n || M->ncols != ncols || !M->ptr || M->ptr[0] != 0) return false;
    for (i64 i = 0; i < n; i++)
        if (M->ptr[i + 1] < M->ptr[i]) return false;
    const i64 nnz = M->ptr[n];
    if (nnz > 0 && (!M->col || !M->val)) return false;
    for (i64 p = 0; p < nnz; p++)
        if (M->col[p] < 0 || M->col[p] >= ncols) return false;
    return true;
}

static bool perm_ok(const int32_t *p, i64 n)
{
    if (!p) return false;
    std::vector<char> seen((size_t)n, 0);
    for (i64 i = 0; i < n; i++) {
        if (p[i] < 0 || p[i] >= n || seen[(size_t)p[i]]) return false;
        seen[(size_t)p[i]] = 1;
    }
    return true;
}

namespace {
// host staging of one triangular factor
struct HTri {
    std::vector<i64> ptr;
    std::vector<int> col, order;
    std::vector<Z> val, dinv;
    std::vector<i64> lvl;
    std::vector<Seg> segs;
    int lanes = 2;
};
}  // namespace

static bool plan_tri(const bbjd_csr *F, i64 n, int lower, HTri &h)
{
    std::vector<int32_t> level((size_t)n);
    const i64 nlev = bbjd_sparse_levels(n, F->ptr, F->col, lower, level.data());
    if (nlev < 0) return false;
    h.ptr.assign((size_t)n + 1, 0);
    h.dinv.assign((size_t)n, Z{0, 0});
    std::vector<char> has_d((size_t)n, 0);
    for (i64 i = 0; i < n; i++) {
        for (i64 p = F->ptr[i]; p < F->ptr[i + 1]; p++) {
            const Z v{F->val[p].real(), F->val[p].imag()};
            if (F->col[p] == i) {
                const double den = v.re * v.re + v.im * v.im;
                if (!(den > 0) || has_d[(size_t)i]) return false;
                h.dinv[(size_t)i] = Z{v.re / den, -v.im / den};
                has_d[(size_t)i] = 1;
            } else {
                h.col.push_back(F->col[p]);
                h.val.push_back(v);
            }
        }
        if (!has_d[(size_t)i]) return false;
        h.ptr[(size_t)i + 1] = (i64)h.col.size();
    }
    // rows by level (counting sort: rows of a level in increasing order)
    h.lvl.assign((size_t)nlev + 1, 0);
    for (i64 i = 0; i < n; i++) h.lvl[(size_t)level[(size_t)i] + 1]++;
    for (i64 l = 0; l < nlev; l++) h.lvl[(size_t)l + 1] += h.lvl[(size_t)l];
    h.order.assign((size_t)n, 0);
    std::vector<i64> fill(h.lvl.begin(), h.lvl.end() - 1);
    for (i64 i = 0; i < n; i++) h.order[(size_t)fill[(size_t)level[(size_t)i]]++] = (int)i;
    h.lanes = lanes_for((i64)h.col.size(), n);
    // a level wider than four passes of the workgroup walker gets a grid
    const i64 wide_rows = 4 * (kNarrowTpb / h.lanes);
    for (i64 l = 0; l < nlev; l++) {
        const i64 rows = h.lvl[(size_t)l + 1] - h.lvl[(size_t)l];
        if (rows >= wide_rows) h.segs.push_back(Seg{1, l, l + 1});
        else if (!h.segs.empty() && !h.segs.back().wide) h.segs.back().l1 = l + 1;
        else h.segs.push_back(Seg{0, l, l + 1});
    }
    return true;
}

namespace {
// offsets of the single device allocation
struct Layout {
    size_t off = 0;
    size_t take(size_t bytes)
    {
        const size_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    }
};
}  // namespace

static void sparse_free_impl(bbjd_sparse_op *op)
{
    if (!op) return;
    if (op->block) {
        if (op->stream) (void)hipStreamSynchronize(op->stream);
        (void)hipFree(op->block);
    }
    delete op;
}

extern "C" {

static bbjd_sparse_op *create_impl(const bbjd_csr *A, const bbjd_csr *B, const bbjd_sparse_lu *lu,
                                   const bbjd_sparse_border *bd, void *stream)
{
    if (!A || A->n <= 0 || A->n > 0x7fffffff) return nullptr;
    const i64 n = A->n;
    if (!csr_ok(A, n, n) || (B && !csr_ok(B, n, n))) return nullptr;
    // pinned rows: R increasing in [0, n), the rest R' in order; the LU is
    // M11's (size m = n - d) and its row permutation is composed with R'
    const int d = bd ? (int)bd->d : 0;
    std::vector<int> rest;
    if (bd) {
        if (!lu || d < 1 || d > 64 || d >= n || !bd->rows || !bd->M21 || !bd->Y || !bd->Sinv) return nullptr;
        for (int j = 0; j < d; j++)
            if (bd->rows[j] < 0 || bd->rows[j] >= n || (j > 0 && bd->rows[j] <= bd->rows[j - 1])) return nullptr;
        rest.reserve((size_t)(n - d));
        for (i64 i = 0, j = 0; i < n; i++) {
            if (j < d && bd->rows[j] == i) { j++; continue; }
            rest.push_back((int)i);
        }
    }
    const i64 m = n - d;
    HTri hl, hu;
    std::vector<int> iperm_r, perm_c;
    if (lu) {
        if (!csr_ok(&lu->L, m, m) || !csr_ok(&lu->U, m, m) || !perm_ok(lu->perm_r, m) || !perm_ok(lu->perm_c, m))
            return nullptr;
        if (!plan_tri(&lu->L, m, 1, hl) || !plan_tri(&lu->U, m, 0, hu)) return nullptr;
        iperm_r.assign((size_t)m, 0);
        perm_c.assign(lu->perm_c, lu->perm_c + m);
        for (i64 i = 0; i < m; i++) iperm_r[(size_t)lu->perm_r[i]] = (int)i;
        if (bd)
            for (i64 i = 0; i < m; i++) iperm_r[(size_t)i] = rest[(size_t)iperm_r[(size_t)i]];
    }
    const i64 nnzA = A->ptr[n], nnzB = B ? B->ptr[n] : 0;
    Layout L;
    const size_t oA_ptr = L.take((size_t)(n + 1) * 8), oA_col = L.take((size_t)nnzA * 4),
                 oA_val = L.take((size_t)nnzA * 16);
    const size_t oB_ptr = B ? L.take((size_t)(n + 1) * 8) : 0, oB_col = B ? L.take((size_t)nnzB * 4) : 0,
                 oB_val = B ? L.take((size_t)nnzB * 16) : 0;
    size_t oT[2][6] = {{0}};
    HTri *ht[2] = {&hl, &hu};
    if (lu)
        for (int t = 0; t < 2; t++) {
            oT[t][0] = L.take(ht[t]->ptr.size() * 8);
            oT[t][1] = L.take(ht[t]->col.size() * 4);
            oT[t][2] = L.take(ht[t]->val.size() * 16);
            oT[t][3] = L.take((size_t)m * 16);
            oT[t][4] = L.take((size_t)m * 4);
            oT[t][5] = L.take(ht[t]->lvl.size() * 8);
        }
    const size_t oIr = lu ? L.take((size_t)m * 4) : 0, oPc = lu ? L.take((size_t)m * 4) : 0;
    const size_t oS = L.take((size_t)n * 16);
    size_t oRows = 0, oRest = 0, oM21 = 0, oY = 0, oSinv = 0, oT2 = 0, oXr = 0, oZ = 0;
    if (bd) {
        oRows = L.take((size_t)d * 4); oRest = L.take((size_t)m * 4);
        oM21 = L.take((size_t)d * m * 16); oY = L.take((size_t)m * d * 16); oSinv = L.take((size_t)d * d * 16);
        oT2 = L.take((size_t)d * 16); oXr = L.take((size_t)d * 16); oZ = L.take((size_t)m * 16);
    }

    bbjd_sparse_op *op = new (std::nothrow) bbjd_sparse_op();
    if (!op) return nullptr;
    op->n = n;
    op->stream = (hipStream_t)stream;
    op->has_B = B != nullptr;
    op->has_lu = lu != nullptr;
    const int held = bbjd_devwide_lock();
    const hipError_t ea = hipMalloc(&op->block, L.off);
    bbjd_devwide_unlock(held);
    if (ea != hipSuccess) { op->block = nullptr; sparse_free_impl(op); return nullptr; }
    op->bytes = L.off;
    char *b = (char *)op->block;
    hipStream_t st = op->stream;
    int rc = 0;
    auto up = [&](size_t off, const void *src, size_t bytes) {
        if (bytes) rc |= hipMemcpyAsync(b + off, src, bytes, hipMemcpyHostToDevice, st) != hipSuccess;
    };
    // values: bbjd_zomplex is (re, im) of double, the layout of Z
    op->A.ptr = (i64 *)(b + oA_ptr); op->A.col = (int *)(b + oA_col); op->A.val = (Z *)(b + oA_val);
    op->A.nnz = nnzA; op->A.lanes = lanes_for(nnzA, n);
    up(oA_ptr, A->ptr, (size_t)(n + 1) * 8); up(oA_col, A->col, (size_t)nnzA * 4); up(oA_val, A->val, (size_t)nnzA * 16);
    if (B) {
        op->B.ptr = (i64 *)(b + oB_ptr); op->B.col = (int *)(b + oB_col); op->B.val = (Z *)(b + oB_val);
        op->B.nnz = nnzB; op->B.lanes = lanes_for(nnzB, n);
        up(oB_ptr, B->ptr, (size_t)(n + 1) * 8); up(oB_col, B->col, (size_t)nnzB * 4);
        up(oB_val, B->val, (size_t)nnzB * 16);
    }
    if (lu) {
        DTri *dt[2] = {&op->L, &op->U};
        for (int t = 0; t < 2; t++) {
            HTri &h = *ht[t];
            DTri &d = *dt[t];
            d.off.ptr = (i64 *)(b + oT[t][0]); d.off.col = (int *)(b + oT[t][1]); d.off.val = (Z *)(b + oT[t][2]);
            d.off.nnz = (i64)h.col.size(); d.off.lanes = h.lanes;
            d.dinv = (Z *)(b + oT[t][3]); d.order = (int *)(b + oT[t][4]); d.lvl = (i64 *)(b + oT[t][5]);
            d.nlev = (i64)h.lvl.size() - 1;
            d.hlvl = h.lvl;
            d.segs = h.segs;
            up(oT[t][0], h.ptr.data(), h.ptr.size() * 8);
            up(oT[t][1], h.col.data(), h.col.size() * 4);
            up(oT[t][2], h.val.data(), h.val.size() * 16);
            up(oT[t][3], h.dinv.data(), (size_t)m * 16);
            up(oT[t][4], h.order.data(), (size_t)m * 4);
            up(oT[t][5], h.lvl.data(), h.lvl.size() * 8);
        }
        op->iperm_r = (int *)(b + oIr);
        op->perm_c = (int *)(b + oPc);
        up(oIr, iperm_r.data(), (size_t)m * 4);
        up(oPc, perm_c.data(), (size_t)m * 4);
    }
    op->s = (Z *)(b + oS);
    if (bd) {
        op->has_border = 1; op->d = d; op->m = m;
        op->rows = (int *)(b + oRows); op->rest = (int *)(b + oRest);
        op->M21 = (Z *)(b + oM21); op->Y = (Z *)(b + oY); op->Sinv = (Z *)(b + oSinv);
        op->t = (Z *)(b + oT2); op->xr = (Z *)(b + oXr); op->z = (Z *)(b + oZ);
        up(oRows, bd->rows, (size_t)d * 4); up(oRest, rest.data(), (size_t)m * 4);
        up(oM21, bd->M21, (size_t)d * m * 16); up(oY, bd->Y, (size_t)m * d * 16);
        up(oSinv, bd->Sinv, (size_t)d * d * 16);
    }
    // the host arrays go out of scope: the copies must have completed
    rc |= hipStreamSynchronize(st) != hipSuccess;
    if (rc) { sparse_free_impl(op); return nullptr; }
    return op;
}

bbjd_sparse_op *bbjd_sparse_op_create(const bbjd_csr *A, const bbjd_csr *B, const bbjd_sparse_lu *lu, void *stream)
{
    return create_impl(A, B, lu, nullptr, stream);
}

bbjd_sparse_op *bbjd_sparse_op_create_bordered(const bbjd_csr *A, const bbjd_csr *B, const bbjd_sparse_lu *lu,
                                               const bbjd_sparse_border *border, void *stream)
{
    if (!border) return nullptr;
    return create_impl(A, B, lu, border, stream);
}

void bbjd_sparse_op_free(bbjd_sparse_op *op)
{
    const int held = bbjd_devwide_lock();
    sparse_free_impl(op);
    bbjd_devwide_unlock(held);
}

size_t bbjd_sparse_op_bytes(const bbjd_sparse_op *op) { return op ? op->bytes : 0; }

int bbjd_sparse_op_levels(const bbjd_sparse_op *op, int64_t *levels_L, int64_t *levels_U, int64_t *launches)
{
    if (!op) return -1;
    if (levels_L) *levels_L = op->has_lu ? op->L.nlev : 0;
    if (levels_U) *levels_U = op->has_lu ? op->U.nlev : 0;
    if (launches)
        *launches = op->has_lu ? (int64_t)(op->L.segs.size() + op->U.segs.size() + (op->has_border ? 5 : 2)) : 1;
    return 0;
}

int bbjd_sparse_matvec_A(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *data)
{
    bbjd_sparse_op *op = (bbjd_sparse_op *)data;
    if (!op || n != op->n) return -1;
    return launch_mv(sp_stream(op), n, op->A, x, y);
}

int bbjd_sparse_matvec_B(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *data)
{
    bbjd_sparse_op *op = (bbjd_sparse_op *)data;
    if (!op || n != op->n) return -1;
    hipStream_t st = sp_stream(op);
    if (op->has_B) return launch_mv(st, n, op->B, x, y);
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n, kTpb)), dim3(kTpb), 0, st, (i64)n, (const Z *)x, (Z *)y);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int bbjd_sparse_precond(bbjd_int n, const bbjd_zomplex *x, bbjd_zomplex *y, void *data)
{
    bbjd_sparse_op *op = (bbjd_sparse_op *)data;
    if (!op || n != op->n) return -1;
    hipStream_t st = sp_stream(op);
    const int g = grid_for(n, kTpb);
    if (!op->has_lu) {
        hipLaunchKernelGGL(k_copy, dim3(g), dim3(kTpb), 0, st, (i64)n, (const Z *)x, (Z *)y);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    if (op->has_border) {
        // s = Pr x[R'] and x_R read first (y may be x), s <- U^-1 L^-1 s,
        // z = Pc s = M11^-1 x[R'], t = M21 z, then the Schur complement
        const i64 m = op->m;
        const int gm = grid_for(m, kTpb);
        hipLaunchKernelGGL(k_gather, dim3(gm), dim3(kTpb), 0, st, m, (const int *)op->iperm_r, (const Z *)x, op->s);
        hipLaunchKernelGGL(k_gather, dim3(1), dim3(64), 0, st, (i64)op->d, (const int *)op->rows, (const Z *)x, op->xr);
        if (hipGetLastError() != hipSuccess) return -1;
        if (launch_tri(st, op->L, op->s) || launch_tri(st, op->U, op->s)) return -1;
        hipLaunchKernelGGL(k_gather, dim3(gm), dim3(kTpb), 0, st, m, (const int *)op->perm_c, (const Z *)op->s, op->z);
        hipLaunchKernelGGL(k_border_dots, dim3(op->d), dim3(kTpb), 0, st, m, (const Z *)op->M21, (const Z *)op->z,
                           op->t);
        hipLaunchKernelGGL(k_border_combine, dim3(gm), dim3(kTpb), 0, st, m, op->d, (const int *)op->rows,
                           (const int *)op->rest, (const Z *)op->Sinv, (const Z *)op->Y, (const Z *)op->t,
                           (const Z *)op->xr, (const Z *)op->z, (Z *)y);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    // s = Pr x, s <- L^-1 s, s <- U^-1 s, y = Pc s
    hipLaunchKernelGGL(k_gather, dim3(g), dim3(kTpb), 0, st, (i64)n, (const int *)op->iperm_r, (const Z *)x, op->s);
    if (hipGetLastError() != hipSuccess) return -1;
    if (launch_tri(st, op->L, op->s) || launch_tri(st, op->U, op->s)) return -1;
    hipLaunchKernelGGL(k_gather, dim3(g), dim3(kTpb), 0, st, (i64)n, (const int *)op->perm_c, (const Z *)op->s, (Z *)y);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
