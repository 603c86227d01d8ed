// qz_dev.hip -- the projected step on the device: batched, one 64-lane
// wavefront per problem, everything in LDS (include/bbjd_gpu.h,
// bbjd_qz_device_*).
//
// The reference solves the k x k projected pencil (MA, MB) with ZGGES,
// selects the eigenvalues matching converged ones plus the nkeep best by the
// target key, moves them to the leading block with ZTGSEN, and takes the
// best unconverged value over all positions as the target
// (src/jd_core.c:319-462), then back-substitutes the Ritz coefficients
// (:851-870).  The engine's test space makes C = nu0 MA + mu0 MB upper
// triangular, so MA x = theta MB x is the standard problem
// M x = x / (nu0 theta + mu0) with M = C^-1 MB: the same right invariant
// subspaces, so an ordered Schur form M = Z T Z^H serves the restart in
// place of ZB, and the Ritz vector is Z y with y from T.  Steps per problem:
//   1. C, its triangularity and diagonal checked; M = C^-1 MB by columns;
//   2. Householder reduction to Hessenberg form, Z accumulated;
//   3. single-shift QR (the trailing 2 x 2 block's eigenvalue nearer its
//      last diagonal entry, an exceptional shift every tenth iteration),
//      deflation by |h(i, i-1)| <= eps (|h(i-1, i-1)| + |h(i, i)|) in the
//      1-norm of the parts, Z accumulated;
//   4. theta_i = (1 - mu0 mu_i) / (nu0 mu_i); the reference's selection;
//      the selected values moved to the leading positions in their order by
//      swaps of adjacent diagonal entries (a rotation that zeroes the
//      difference of the two after the exchange);
//   5. the target by the reference's scan of all positions; the Ritz
//      coefficients x = Z [y; 1; 0] from (T - t_pp I) y = -T(:, p).
// A problem the host would decide differently is flagged instead (info[0]):
// C not triangular or singular, QR not converged, a choice that the thetas'
// error bounds could change (as in the host's harmonic step: a converged
// match, the restart's selection, the target, or a target off by more than
// 10 tol), a tie for the target (up to rounding) or no unconverged
// candidate; the engine then runs the host step for that iteration.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "../../include/bbjd_gpu.h"
#include "engine.h"

typedef long long i64;

namespace {

constexpr int QK = 64;          // largest k
constexpr int QL = QK + 1;      // leading dimension in LDS (bank spread)
constexpr int QT = 64;          // one wavefront

struct Z2 { double re, im; };
__device__ __forceinline__ Z2 zmk(double a, double b) { return Z2{a, b}; }
__device__ __forceinline__ Z2 zadd(Z2 a, Z2 b) { return zmk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ Z2 zsub(Z2 a, Z2 b) { return zmk(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ Z2 zmul(Z2 a, Z2 b) { return zmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__device__ __forceinline__ Z2 zconj(Z2 a) { return zmk(a.re, -a.im); }
__device__ __forceinline__ Z2 zscal(double s, Z2 a) { return zmk(s * a.re, s * a.im); }
__device__ __forceinline__ double zabs(Z2 a) { return hypot(a.re, a.im); }
__device__ __forceinline__ double zabs1(Z2 a) { return fabs(a.re) + fabs(a.im); }
__device__ __forceinline__ Z2 zdiv(Z2 a, Z2 b)
{
    // Smith's algorithm
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + r * b.im;
        return zmk((a.re + r * a.im) / d, (a.im - r * a.re) / d);
    }
    const double r = b.re / b.im, d = b.im + r * b.re;
    return zmk((r * a.re + a.im) / d, (r * a.im - a.re) / d);
}
// principal square root: t = sqrt((|z| + |x|) / 2) formed without
// cancellation, the other part as y / (2 t)
__device__ __forceinline__ Z2 zsqrt(Z2 z)
{
    const double r = zabs(z);
    if (r == 0.0) return zmk(0.0, 0.0);
    const double t = sqrt(0.5 * (r + fabs(z.re)));
    if (z.re >= 0.0) return zmk(t, z.im / (2.0 * t));
    return zmk(fabs(z.im) / (2.0 * t), copysign(t, z.im));
}

// rotation with real c >= 0: [c s; -conj(s) c] [x; y] = [r; 0]
__device__ __forceinline__ void zrot(Z2 x, Z2 y, double &c, Z2 &s, Z2 &r)
{
    const double ax = zabs(x), ay = zabs(y);
    if (ay == 0.0) { c = 1.0; s = zmk(0.0, 0.0); r = x; return; }
    if (ax == 0.0) { c = 0.0; s = zscal(1.0 / ay, zconj(y)); r = zmk(ay, 0.0); return; }
    const double rho = hypot(ax, ay);
    const Z2 ph = zscal(1.0 / ax, x);
    c = ax / rho;
    s = zscal(1.0 / rho, zmul(ph, zconj(y)));
    r = zscal(rho, ph);
}

__device__ __forceinline__ double key_of(int tgt, Z2 lam, Z2 sigma)
{
    switch (tgt) {
    case BBJD_TARGET_LARGEST_MAG:   return -zabs(lam);
    case BBJD_TARGET_SMALLEST_MAG:  return zabs(lam);
    case BBJD_TARGET_LARGEST_REAL:  return -lam.re;
    case BBJD_TARGET_SMALLEST_REAL: return lam.re;
    case BBJD_TARGET_LARGEST_IMAG:  return -lam.im;
    case BBJD_TARGET_SMALLEST_IMAG: return lam.im;
    default:                        return zabs(zsub(lam, sigma));
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

struct Qs {                     // LDS of one problem
    Z2 T[QK * QL];              // T(i, j) at i + j QL
    Z2 Zm[QK * QL];             // Z(i, j) at i + j QL
    Z2 v[QK];                   // reflector / coefficients
    Z2 th[QK];                  // theta per position
    Z2 al[QK], be[QK];
    double eb[QK];              // error bound of theta per position
    int sel[QK];
    int flag;
};

// rows i, i + 1 of T, columns j0 .. k - 1 (lane = column)
__device__ __forceinline__ void rot_rows(Z2 *T, int i, double c, Z2 s, int j0, int k, int lane)
{
    for (int j = j0 + lane; j < k; j += QT) {
        const Z2 a = T[i + j * QL], b = T[i + 1 + j * QL];
        T[i + j * QL] = zadd(zscal(c, a), zmul(s, b));
        T[i + 1 + j * QL] = zsub(zscal(c, b), zmul(zconj(s), a));
    }
}
// columns i, i + 1 of X, rows 0 .. r1 - 1: X <- X G^H (lane = row)
__device__ __forceinline__ void rot_cols(Z2 *X, int i, double c, Z2 s, int r1, int lane)
{
    for (int r = lane; r < r1; r += QT) {
        const Z2 a = X[r + i * QL], b = X[r + (i + 1) * QL];
        X[r + i * QL] = zadd(zscal(c, a), zmul(zconj(s), b));
        X[r + (i + 1) * QL] = zsub(zscal(c, b), zmul(s, a));
    }
}

}  // namespace

__global__ __launch_bounds__(QT) void k_qz_step(const bbjd_qz_problem *__restrict__ probs)
{
    __shared__ Qs sm;
    const bbjd_qz_problem P = probs[blockIdx.x];
    const int lane = threadIdx.x;
    const int k = P.k;
    int *info = P.info;
    const double eps = 2.220446049250313e-16;
    if (k < 1 || k > QK) {
        if (lane == 0) { info[0] = BBJD_QZD_UNSUPPORTED; info[1] = 0; info[2] = 0; info[3] = 0; }
        return;
    }
    const Z2 nu0 = zmk(P.nu0_re, P.nu0_im), mu0 = zmk(P.mu0_re, P.mu0_im), sigma = zmk(P.sigma_re, P.sigma_im);
    const Z2 *MA = reinterpret_cast<const Z2 *>(P.MA);
    const Z2 *MB = reinterpret_cast<const Z2 *>(P.MB);
    Z2 *T = sm.T, *Zm = sm.Zm;
    // ---- 1. C = nu0 MA + mu0 MB in T, MB in Zm; checks ----
    double cmax = 0, lmax = 0, dmin = INFINITY;
    for (int j = 0; j < k; j++) {
        if (lane < k) {
            const Z2 a = MA[lane + (i64)j * P.ldm], b = MB[lane + (i64)j * P.ldm];
            const Z2 c = zadd(zmul(nu0, a), zmul(mu0, b));
            T[lane + j * QL] = c;
            Zm[lane + j * QL] = b;
            const double ac = zabs(c);
            cmax = fmax(cmax, ac);
            if (lane > j) lmax = fmax(lmax, ac);
            if (lane == j) dmin = fmin(dmin, ac);
        }
    }
    for (int m = 32; m >= 1; m >>= 1) {
        cmax = fmax(cmax, __shfl_xor(cmax, m, 64));
        lmax = fmax(lmax, __shfl_xor(lmax, m, 64));
        dmin = fmin(dmin, __shfl_xor(dmin, m, 64));
    }
    __syncthreads();
    if (!(cmax > 0) || lmax > 1e-12 * cmax || !(dmin > 1e-300) || zabs(nu0) == 0.0) {
        if (lane == 0) { info[0] = BBJD_QZD_NOT_TRIANGULAR; info[1] = 0; info[2] = 0; info[3] = 0; }
        return;
    }
    // M = C^-1 MB: column j (lane = j) by back-substitution over the rows
    if (lane < k) {
        const int j = lane;
        for (int i = k - 1; i >= 0; i--) {
            Z2 acc = Zm[i + j * QL];
            for (int q = i + 1; q < k; q++) acc = zsub(acc, zmul(T[i + q * QL], Zm[q + j * QL]));
            Zm[i + j * QL] = zdiv(acc, T[i + i * QL]);
        }
    }
    __syncthreads();
    // T <- M, Zm <- I; ||M||_F
    double fro = 0;
    for (int j = 0; j < k; j++) {
        if (lane < k) {
            const Z2 m = Zm[lane + j * QL];
            fro += m.re * m.re + m.im * m.im;
        }
    }
    __syncthreads();
    for (int j = 0; j < k; j++)
        if (lane < k) {
            T[lane + j * QL] = Zm[lane + j * QL];
            Zm[lane + j * QL] = zmk(lane == j ? 1.0 : 0.0, 0.0);
        }
    fro = sqrt(wave_sum(fro));
    __syncthreads();
    if (!isfinite(fro)) {
        if (lane == 0) { info[0] = BBJD_QZD_ACCURACY; info[1] = 0; info[2] = 0; info[3] = 0; }
        return;
    }
    // ---- 2. Hessenberg reduction, Z accumulated ----
    for (int j = 0; j + 2 < k; j++) {
        double nx = 0;
        if (lane > j && lane < k) {
            const Z2 x = T[lane + j * QL];
            nx = x.re * x.re + x.im * x.im;
        }
        nx = wave_sum(nx);
        const double alpha = sqrt(nx);
        const Z2 x0 = T[j + 1 + j * QL];
        const double ax0 = zabs(x0);
        if (alpha == 0.0) continue;               // uniform over the wave
        const Z2 ph = ax0 > 0 ? zscal(1.0 / ax0, x0) : zmk(1.0, 0.0);
        // v = x + ph alpha e1, ||v||^2 = 2 alpha (alpha + |x0|)
        const double vn = sqrt(2.0 * alpha * (alpha + ax0));
        if (lane > j && lane < k) {
            Z2 x = T[lane + j * QL];
            if (lane == j + 1) x = zadd(x, zscal(alpha, ph));
            sm.v[lane] = zscal(1.0 / vn, x);
        }
        __syncthreads();
        // T(j+1:, :) -= 2 v (v^H T(j+1:, :)), lane = column
        for (int c = lane; c < k; c += QT) {
            Z2 w = zmk(0, 0);
            for (int i = j + 1; i < k; i++) w = zadd(w, zmul(zconj(sm.v[i]), T[i + c * QL]));
            w = zscal(2.0, w);
            for (int i = j + 1; i < k; i++) T[i + c * QL] = zsub(T[i + c * QL], zmul(sm.v[i], w));
        }
        __syncthreads();
        // T(:, j+1:) -= 2 (T(:, j+1:) v) v^H and Z likewise, lane = row
        for (int r = lane; r < k; r += QT) {
            Z2 u = zmk(0, 0), uz = zmk(0, 0);
            for (int i = j + 1; i < k; i++) {
                u = zadd(u, zmul(T[r + i * QL], sm.v[i]));
                uz = zadd(uz, zmul(Zm[r + i * QL], sm.v[i]));
            }
            u = zscal(2.0, u);
            uz = zscal(2.0, uz);
            for (int i = j + 1; i < k; i++) {
                T[r + i * QL] = zsub(T[r + i * QL], zmul(u, zconj(sm.v[i])));
                Zm[r + i * QL] = zsub(Zm[r + i * QL], zmul(uz, zconj(sm.v[i])));
            }
        }
        __syncthreads();
        if (lane > j + 1 && lane < k) T[lane + j * QL] = zmk(0, 0);
        __syncthreads();
    }
    // ---- 3. shifted QR to the Schur form ----
    int ihi = k - 1, its = 0, total = 0, ok = 1;
    while (ihi >= 0) {
        // deflation: the largest i in [1, ihi] with a negligible subdiagonal
        bool small = false;
        if (lane >= 1 && lane <= ihi) {
            const Z2 h = T[lane + (lane - 1) * QL];
            small = zabs1(h) <= eps * (zabs1(T[lane - 1 + (lane - 1) * QL]) + zabs1(T[lane + lane * QL]));
        }
        const unsigned long long m = __ballot(small);
        int l = 0;
        if (m) l = 63 - __clzll(m);
        __syncthreads();
        if (l > 0 && lane == 0) T[l + (l - 1) * QL] = zmk(0, 0);
        __syncthreads();
        if (l == ihi) { ihi--; its = 0; continue; }
        its++;
        total++;
        if (its > 60 || total > 60 * k) { ok = 0; break; }
        Z2 mu;
        const Z2 d = T[ihi + ihi * QL];
        if (its % 10 == 0) {
            mu = zadd(d, zmk(0.75 * fabs(T[ihi + (ihi - 1) * QL].re), 0.0));
        } else {
            const Z2 a = T[ihi - 1 + (ihi - 1) * QL], b = T[ihi - 1 + ihi * QL], c = T[ihi + (ihi - 1) * QL];
            const Z2 h = zscal(0.5, zsub(a, d));
            const Z2 disc = zsqrt(zadd(zmul(h, h), zmul(b, c)));
            const Z2 m1 = zadd(zadd(d, h), disc), m2 = zsub(zadd(d, h), disc);
            mu = zabs(zsub(m1, d)) <= zabs(zsub(m2, d)) ? m1 : m2;
        }
        for (int kk = l; kk < ihi; kk++) {
            Z2 x, y;
            if (kk == l) { x = zsub(T[l + l * QL], mu); y = T[l + 1 + l * QL]; }
            else { x = T[kk + (kk - 1) * QL]; y = T[kk + 1 + (kk - 1) * QL]; }
            double c;
            Z2 s, r;
            zrot(x, y, c, s, r);
            __syncthreads();
            rot_rows(T, kk, c, s, kk == l ? kk : kk - 1, k, lane);
            __syncthreads();
            if (kk > l && lane == 0) T[kk + 1 + (kk - 1) * QL] = zmk(0, 0);
            rot_cols(T, kk, c, s, kk + 3 < ihi + 1 ? kk + 3 : ihi + 1, lane);
            rot_cols(Zm, kk, c, s, k, lane);
            __syncthreads();
        }
    }
    if (!ok) {
        if (lane == 0) { info[0] = BBJD_QZD_NO_CONVERGENCE; info[1] = 0; info[2] = 0; info[3] = 0; }
        return;
    }
    // ---- 4. theta, error bounds, selection, reordering ----
    // The QR gives each mu = t_ii with an absolute error of about
    // eps ||M||; theta = 1 / (nu0 mu) - mu0 / nu0 then moves by
    // |d mu| / (|nu0| |mu|^2).  A choice that these bounds could change
    // (a converged match within 100 tol, the restart's selection, the
    // target) is flagged for the host, whose generalized QZ decides.
    const double dmu = 4.0 * eps * fro;
    const Z2 *conv = reinterpret_cast<const Z2 *>(P.conv);
    const double tol = P.tol, lim = 100.0 * tol;
    auto diag_values = [&]() {
        if (lane < k) {
            const Z2 m = T[lane + lane * QL];
            const double am = zabs(m);
            sm.al[lane] = zsub(zmk(1.0, 0.0), zmul(mu0, m));
            sm.be[lane] = zmul(nu0, m);
            sm.eb[lane] = am * zabs(nu0) > 1e-300 ? dmu / (zabs(nu0) * am * am) : INFINITY;
        }
    };
    auto conv_dist = [&](Z2 lam) {
        double m = INFINITY;
        for (int c2 = 0; c2 < P.nconv; c2++) m = fmin(m, zabs(zsub(lam, conv[c2])));
        return m;
    };
    diag_values();
    __syncthreads();
    if (lane == 0) {
        sm.flag = 0;
        for (int i = 0; i < k; i++) {
            sm.sel[i] = 0;
            if (zabs(sm.be[i]) < 1e-300) continue;
            const Z2 lam = zdiv(sm.al[i], sm.be[i]);
            sm.th[i] = lam;
            const double m = conv_dist(lam);
            if (m < lim) sm.sel[i] = 1;
            // a restart keeps the selected columns: their membership must be certain
            if (P.keep > 0 && fabs(m - lim) <= sm.eb[i]) sm.flag = 1;
        }
        double kmax = -INFINITY, emax = 0;
        for (int pick = 0; pick < P.nkeep; pick++) {
            int best = -1;
            double bv = INFINITY;
            for (int i = 0; i < k; i++) {
                if (sm.sel[i] || zabs(sm.be[i]) < 1e-300) continue;
                const double v = key_of(P.target, sm.th[i], sigma);
                if (v < bv) { bv = v; best = i; }
            }
            if (best < 0) break;
            sm.sel[best] = 1;
            if (bv + sm.eb[best] > kmax + emax) { kmax = bv; emax = sm.eb[best]; }
        }
        if (P.keep > 0 && kmax > -INFINITY)
            for (int i = 0; i < k; i++)
                if (!sm.sel[i] && zabs(sm.be[i]) >= 1e-300 &&
                    !(key_of(P.target, sm.th[i], sigma) - sm.eb[i] > kmax + emax)) sm.flag = 1;
    }
    __syncthreads();
    if (sm.flag) {
        if (lane == 0) { info[0] = BBJD_QZD_ACCURACY; info[1] = 0; info[2] = 0; info[3] = 0; }
        return;
    }
    int pos = 0;
    for (int i = 0; i < k; i++) {
        if (!sm.sel[i]) continue;
        for (int j = i - 1; j >= pos; j--) {
            const Z2 t11 = T[j + j * QL], t22 = T[j + 1 + (j + 1) * QL], t12 = T[j + (j + 1) * QL];
            double c;
            Z2 s, r;
            zrot(t12, zsub(t22, t11), c, s, r);
            __syncthreads();
            rot_rows(T, j, c, s, j, k, lane);
            __syncthreads();
            rot_cols(T, j, c, s, j + 2, lane);
            rot_cols(Zm, j, c, s, k, lane);
            __syncthreads();
            if (lane == 0) {
                T[j + 1 + j * QL] = zmk(0, 0);
                T[j + j * QL] = t22;
                T[j + 1 + (j + 1) * QL] = t11;
            }
            __syncthreads();
        }
        pos++;
    }
    // ---- 5. target and Ritz coefficients ----
    diag_values();
    __syncthreads();
    int p = 0, status = 0;
    {
        double bv = INFINITY, bv2 = INFINITY;
        for (int i = 0; i < k; i++) {
            if (zabs(sm.be[i]) < 1e-300) continue;
            const Z2 lam = zdiv(sm.al[i], sm.be[i]);
            if (conv_dist(lam) < lim) continue;
            const double v = key_of(P.target, lam, sigma);
            if (v < bv) { bv2 = bv; bv = v; p = i; }
            else if (v < bv2) bv2 = v;
        }
        // a tie or no candidate goes to the host's ordered Schur form
        // whatever the bounds; otherwise the choice must hold for every
        // value within its bound
        if (bv == INFINITY) status = BBJD_QZD_AMBIGUOUS;
        else if (bv2 < INFINITY && bv2 - bv <= 1e-10 * fmax(fmax(fabs(bv), fabs(bv2)), 1e-14)) status = BBJD_QZD_AMBIGUOUS;
        else {
            const double ep = sm.eb[p];
            if (!(ep <= 10.0 * tol) || conv_dist(zdiv(sm.al[p], sm.be[p])) - ep < lim) status = BBJD_QZD_ACCURACY;
            for (int i = 0; i < k && status == 0; i++) {
                if (i == p || zabs(sm.be[i]) < 1e-300) continue;
                const Z2 lam = zdiv(sm.al[i], sm.be[i]);
                if (conv_dist(lam) + sm.eb[i] < lim) continue;
                if (!(key_of(P.target, lam, sigma) - sm.eb[i] > bv + ep)) status = BBJD_QZD_ACCURACY;
            }
        }
    }
    if (lane == 0) {
        // (T(0:p, 0:p) - t_pp I) y = -T(0:p, p), y_p = 1
        const Z2 tpp = T[p + p * QL];
        sm.v[p] = zmk(1.0, 0.0);
        for (int i = p - 1; i >= 0; i--) {
            Z2 acc = T[i + p * QL];
            for (int j = i + 1; j < p; j++) acc = zadd(acc, zmul(T[i + j * QL], sm.v[j]));
            const Z2 dd = zsub(T[i + i * QL], tpp);
            sm.v[i] = zabs(dd) > 1e-300 ? zscal(-1.0, zdiv(acc, dd)) : zmk(0, 0);
        }
    }
    __syncthreads();
    Z2 *xo = reinterpret_cast<Z2 *>(P.x), *ao = reinterpret_cast<Z2 *>(P.alpha), *bo = reinterpret_cast<Z2 *>(P.beta);
    if (lane < k) {
        Z2 acc = zmk(0, 0);
        for (int j = 0; j <= p; j++) acc = zadd(acc, zmul(Zm[lane + j * QL], sm.v[j]));
        if (xo) xo[lane] = acc;
        if (ao) ao[lane] = sm.al[lane];
        if (bo) bo[lane] = sm.be[lane];
    }
    // the leading keep Schur vectors, row-major (zr[j keep + c] = Z(j, c)):
    // the restart's rotation
    if (P.zr && P.keep > 0) {
        Z2 *zr = reinterpret_cast<Z2 *>(P.zr);
        for (int c = 0; c < P.keep && c < k; c++)
            if (lane < k) zr[(i64)lane * P.keep + c] = Zm[lane + c * QL];
    }
    if (P.zfull && lane < k)
        for (int c = 0; c < k; c++) reinterpret_cast<Z2 *>(P.zfull)[lane + (i64)c * k] = Zm[lane + c * QL];
    if (P.tfull && lane < k)
        for (int c = 0; c < k; c++) reinterpret_cast<Z2 *>(P.tfull)[lane + (i64)c * k] = T[lane + c * QL];
    if (lane == 0) {
        int nsel = 0;
        for (int i = 0; i < k; i++) nsel += sm.sel[i];
        info[0] = status;
        info[1] = p;
        info[2] = nsel;
        info[3] = total;
    }
}

extern "C" {

int bbjd_qz_device_batch(const bbjd_qz_problem *probs, int nprob, void *stream)
{
    if (!probs || nprob <= 0) return -1;
    hipLaunchKernelGGL(k_qz_step, dim3((unsigned)nprob), dim3(QT), 0, (hipStream_t)stream, probs);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"

extern "C" int bbjd_qz_device_host(bbjd_int k, const bbjd_zomplex *MA, const bbjd_zomplex *MB, bbjd_int ldm,
                                   bbjd_zomplex nu0, bbjd_zomplex mu0, bbjd_zomplex sigma, int target, bbjd_int nkeep,
                                   const bbjd_zomplex *conv, bbjd_int nconv, double tol, bbjd_zomplex *alpha,
                                   bbjd_zomplex *beta, bbjd_zomplex *x, bbjd_zomplex *z, bbjd_zomplex *t, int *info)
{
    if (k < 1 || k > QK || ldm < k || nconv < 0 || !MA || !MB || !info || (nconv > 0 && !conv)) return -1;
    const size_t mb = (size_t)ldm * k * 16, vb = (size_t)k * 16, kk = (size_t)k * k * 16;
    const size_t off_mb = mb, off_conv = 2 * mb, off_a = off_conv + (size_t)(nconv > 0 ? nconv : 1) * 16,
                 off_b = off_a + vb, off_x = off_b + vb, off_z = off_x + vb, off_t = off_z + kk, off_i = off_t + kk,
                 off_p = (off_i + 64 + 255) / 256 * 256, total = off_p + sizeof(bbjd_qz_problem);
    char *d = nullptr;
    if (hipMalloc((void **)&d, total) != hipSuccess) return -1;
    bbjd_qz_problem P;
    memset(&P, 0, sizeof(P));
    P.k = (int)k; P.ldm = (int)ldm;
    P.MA = (const bbjd_zomplex *)d; P.MB = (const bbjd_zomplex *)(d + off_mb);
    P.nu0_re = nu0.real(); P.nu0_im = nu0.imag(); P.mu0_re = mu0.real(); P.mu0_im = mu0.imag();
    P.sigma_re = sigma.real(); P.sigma_im = sigma.imag();
    P.target = target; P.nkeep = (int)nkeep; P.nconv = (int)nconv; P.keep = 0;
    P.conv = (const bbjd_zomplex *)(d + off_conv);
    P.tol = tol;
    P.alpha = (bbjd_zomplex *)(d + off_a); P.beta = (bbjd_zomplex *)(d + off_b); P.x = (bbjd_zomplex *)(d + off_x);
    P.zfull = (bbjd_zomplex *)(d + off_z); P.tfull = (bbjd_zomplex *)(d + off_t);
    P.info = (int *)(d + off_i);
    int rc = 0;
    hipStream_t st = nullptr;
    rc |= hipMemcpyAsync(d, MA, mb, hipMemcpyHostToDevice, st) != hipSuccess;
    rc |= hipMemcpyAsync(d + off_mb, MB, mb, hipMemcpyHostToDevice, st) != hipSuccess;
    if (nconv > 0) rc |= hipMemcpyAsync(d + off_conv, conv, (size_t)nconv * 16, hipMemcpyHostToDevice, st) != hipSuccess;
    rc |= hipMemcpyAsync(d + off_p, &P, sizeof(P), hipMemcpyHostToDevice, st) != hipSuccess;
    if (!rc) rc |= bbjd_qz_device_batch((const bbjd_qz_problem *)(d + off_p), 1, st);
    auto down = [&](void *dst, size_t off, size_t b) {
        if (dst) rc |= hipMemcpyAsync(dst, d + off, b, hipMemcpyDeviceToHost, st) != hipSuccess;
    };
    down(alpha, off_a, vb); down(beta, off_b, vb); down(x, off_x, vb); down(z, off_z, kk); down(t, off_t, kk);
    down(info, off_i, 4 * sizeof(int));
    rc |= hipStreamSynchronize(st) != hipSuccess;
    (void)hipFree(d);
    return rc ? -1 : 0;
}
