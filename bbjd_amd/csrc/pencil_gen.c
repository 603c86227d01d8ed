/*
 * pencil_gen.c -- seeded generator of block-banded pencils with dense tau
 * rows (layout of include/bbjd_ops.h).
 *
 * Every entry is complex N(0,1) in the CN(0,1) sense (real and imaginary
 * parts independent N(0, 1/2)), drawn by Box-Muller from two 53-bit
 * uniforms.  The uniforms are the splitmix64 finaliser applied to a counter
 * (seed, stream, 2 index) and (seed, stream, 2 index + 1), so an entry's
 * value depends only on its position: generation order and thread count do
 * not change the pencil.  log/sqrt/cos/sin come from libm.
 */
#include "../../include/bbjd_ops.h"

#include <complex.h>
#include <math.h>
#include <string.h>

static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline double unit53(uint64_t h) { return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

static inline double complex cn01(uint64_t base, uint64_t idx)
{
    const double u1 = unit53(mix64(base + 2 * idx));
    const double u2 = unit53(mix64(base + 2 * idx + 1));
    const double rad = sqrt(-log(u1));
    const double th = 6.283185307179586476925 * u2;
    return rad * cos(th) + I * (rad * sin(th));
}

int bbjd_pencil_generate(bbjd_int n, int bs, int hw, int ntau, uint64_t seed, double diag_shift,
                         bbjd_zomplex *band, bbjd_zomplex *tau, bbjd_zomplex *bdiag)
{
    if (n <= 0 || bs <= 0 || hw < 0 || ntau < 0 || ntau > n || n % bs) return -1;
    const bbjd_int nb = n / bs, n1 = n - ntau;
    const int nd = 2 * hw + 1;
    const uint64_t base_band = mix64(seed ^ mix64(1)), base_tau = mix64(seed ^ mix64(2));
    double complex *B = (double complex *)band, *T = (double complex *)tau, *D = (double complex *)bdiag;
#pragma omp parallel for schedule(static)
    for (bbjd_int bi = 0; bi < nb; bi++) {
        for (int d = 0; d < nd; d++) {
            const bbjd_int bj = bi + d - hw;
            double complex *blk = B + ((size_t)bi * nd + d) * bs * bs;
            for (int c = 0; c < bs; c++)
                for (int r = 0; r < bs; r++) {
                    const size_t idx = ((size_t)bi * nd + d) * bs * bs + (size_t)c * bs + r;
                    const bbjd_int row = bi * bs + r;
                    double complex v = 0.0;
                    if (bj >= 0 && bj < nb && row < n1) {
                        v = cn01(base_band, idx);
                        if (d == hw && r == c) v += diag_shift;
                    }
                    blk[(size_t)c * bs + r] = v;
                }
        }
    }
#pragma omp parallel for schedule(static)
    for (bbjd_int j = 0; j < n; j++)
        for (int t = 0; t < ntau; t++) T[(size_t)t * n + j] = cn01(base_tau, (size_t)t * n + j);
    for (bbjd_int i = 0; i < n; i++) D[i] = i < n1 ? 1.0 : 0.0;
    return 0;
}
