/* ops_dev.h -- launchers of ops.hip (device operators), plain C ABI. */
#ifndef BBJD_OPS_DEV_H
#define BBJD_OPS_DEV_H
#include <stdint.h>
#include <hip/hip_runtime_api.h>
#ifdef __cplusplus
extern "C" {
#endif
int bbjd_o_band_mv(hipStream_t s, int64_t n1, int64_t nb, int bs, int hw, const void *band, const void *x, void *y,
                   int single);
/* y_A = A x (band rows + the ntau dense rows from per-block partials in
 * part, nblk x ntau) and y_B = diag(dg) x in two launches; -3 when ntau > 16
 * or part is too small */
int bbjd_o_band_ab(hipStream_t s, int64_t n, int64_t n1, int64_t nb, int bs, int hw, const void *band,
                   const void *dg, int ntau, const void *tauc, const void *x, void *ya, void *yb, void *part,
                   int64_t part_cap, int single);
/* bbjd_o_band_ab plus r = (are + i aim) y_A + (bre + i bim) y_B and ||r|| ->
 * *rnorm (device double), with per-block |r|^2 partials in rpart
 * (rpart_cap doubles); FP64 only, -3 when ntau > 15 */
int bbjd_o_band_abr(hipStream_t s, int64_t n, int64_t n1, int64_t nb, int bs, int hw, const void *band,
                    const void *dg, int ntau, const void *tauc, const void *x, void *ya, void *yb, void *part,
                    int64_t part_cap, int single, double are, double aim, double bre, double bim, void *r,
                    void *rpart, int64_t rpart_cap, double *rnorm);
int bbjd_o_diag_mv(hipStream_t s, int64_t n, const void *dg, const void *x, void *y, int single);
int bbjd_o_bj_apply(hipStream_t s, int64_t n, int bs, const void *inv, const void *x, void *y, int single);
int bbjd_o_bcr_fwd(hipStream_t st, int s, int64_t M, int64_t h, int64_t ne, const void *AG, void *b, int nrhs,
                   int64_t vs);
/* level 0 of the forward reduction with the right-hand sides read from x0
 * (x1) (n1 rows each): writes all N s rows of b */
int bbjd_o_bcr_fwd0(hipStream_t st, int s, int64_t N, int64_t n1, const void *AG, const void *x0, const void *x1,
                    void *b, int nrhs, int64_t vs);
int bbjd_o_bcr_bwd(hipStream_t st, int s, int64_t M, int64_t h, int64_t no, const void *DPQ, const void *b, void *x,
                   int nrhs, int64_t vs);
/* G (2..4) consecutive levels in one launch, for s in {4, 8, 16}; -3 for
 * other shapes.  Backward levels l + G - 1 .. l (dpq[g]: level l + g's odd
 * factor blocks, bl[g]: the buffer holding its right-hand sides; with tauc
 * and l = 0 also per-block partial dots conj(tauc[:, t]) . x over the rows
 * below n1 into dpart, their block count into *nblk_out, 0 when not made); forward
 * levels l .. l + G - 1 (ag[g]: level l + g's even factor blocks) from the
 * caller's x0 (x1) (l = 0) or from bin into bout (another buffer), writing
 * every row at a multiple of 2^l. */
int bbjd_o_bcr_bwd_multi(hipStream_t st, int s, int64_t N, int l, int G, const void *const *dpq,
                         const void *const *bl, void *x, int nrhs, int64_t vs, const void *tauc, int ntau,
                         int64_t n1, int64_t ldt, void *dpart, int64_t dpart_cap, int64_t *nblk_out);
/* in_c64: x0 is complex64 (one right-hand side, l = 0, s = 4 or 8 with the
 * prefetching kernels; else -3 before any launch) */
int bbjd_o_bcr_fwd_multi(hipStream_t st, int s, int64_t N, int64_t n1, int l, int G, const void *const *ag,
                         const void *x0, const void *x1, const void *bin, void *bout, int nrhs, int64_t vs,
                         int in_c64);
/* first level of the one-workgroup tail for this shape (>= l0, where the
 * LDS-resident tail fits; l0 when the global tail is used) */
int bbjd_o_bcr_tail_level(int s, int64_t N, int l0, int L, int nrhs);
int bbjd_o_bcr_tail(hipStream_t st, int s, int64_t N, int l0, int L, const int64_t *off, const void *F, void *b,
                    void *x, int nrhs, int64_t vs);
/* the border's y2 = Sinv (b2 - dots); with dpart (nblk x ntau x nrhs
 * partials from bbjd_o_bcr_bwd_multi) the dots are summed from it first;
 * c64 (one right-hand side): b2_0 and ytail0 are complex64 */
int bbjd_o_border_small(hipStream_t st, int ntau, const void *b2_0, const void *b2_1, const void *dots0,
                        const void *dots1, const void *Sinv, void *y2, void *ytail0, void *ytail1, int nrhs,
                        const void *dpart, int64_t nblk, int c64);
/* c64 (one right-hand side): y0 is complex64 */
int bbjd_o_border_update(hipStream_t st, int64_t n1, int ntau, const void *xb, int64_t vs, const void *Y,
                         const void *y2, void *y0, void *y1, int nrhs, int c64);
#ifdef __cplusplus
}
#endif
#endif
