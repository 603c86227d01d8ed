/*
 * engine.h -- internal interface between the C host driver and the HIP
 * kernel launchers (kernels.hip, ops.hip).  Plain C ABI: device arrays are
 * passed as void* / typed pointers, complex128 has the layout of double[2],
 * complex64 that of float[2].
 */
#ifndef BBJD_ENGINE_H
#define BBJD_ENGINE_H

#include <stdint.h>
#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BBJD_TPB 256              /* threads per block of streaming kernels */
#define BBJD_TILE_MIN_ROWS 64     /* smallest row tile of any streaming kernel:
                                     per-block norm partials are sized for it */
#define BBJD_PROJ_MIN_ROWS 256    /* smallest row chunk of the projection:
                                     projection partials are sized for it */
#define BBJD_MAXK 4096            /* largest column count of one projection */

/* Scalar slots in the device scalar array (double). */
enum {
    SL_NB = 0,       /* norm before the first CGS pass                  */
    SL_NA,           /* norm after the last CGS pass                    */
    SL_U,            /* ||V x|| of the Ritz vector                      */
    SL_R,            /* ||theta_b A u - theta_a B u||                   */
    SL_V,            /* norm of the new search column                   */
    SL_W,            /* norm of the new test column                     */
    SL_Q,            /* norm of a deflation column                      */
    SL_G,            /* GMRES vector norms                              */
    SL_G2,
    SL_T0, SL_T1,    /* scratch                                         */
    SL_COUNT = 16
};
/* Int flag slots.  FL_RESCALE: the fused CGS2 could not normalise its result
 * in the second pass and left it to the block that finishes its norm. */
enum { FL_PASS2 = 0, FL_ONE = 1, FL_DGKS_COUNT = 2, FL_RESCALE = 3, FL_COUNT = 8 };

/* Counters of the last-block-done reductions: [0, BBJD_CNT_UPD) one per
 * column tile of a projection, then one for updates, one for norms, one for
 * the middle pass of the fused CGS2. */
#define BBJD_CNT_UPD 1024
#define BBJD_CNT_VEC 1025
#define BBJD_CNT_MID 1026
#define BBJD_CNT_SLOTS 1032

/* Per-workspace scratch for the two-stage reductions. */
typedef struct {
    void   *part;        /* partial sums: tiles x ldp complex (<= 3 vectors) */
    double *pnorm;       /* partial norms per tile                          */
    void   *h;           /* coefficient vector, first pass                 */
    void   *h2;          /* second pass                                    */
    void   *hsum;        /* h + h2 (GMRES Hessenberg column)               */
    double *scal;        /* SL_COUNT doubles                                */
    int    *flag;        /* FL_COUNT ints                                   */
    void   *coef;        /* uploaded small vectors                          */
    void   *coef2;
    int    *cnt;         /* BBJD_CNT_SLOTS zero-initialised counters        */
    int64_t part_cap;    /* complex entries available in part               */
} bbjd_red;

/* ---- complex128 streaming kernels (kernels.hip) ---------------------- */

/* CGS2 with the DGKS test (2 passes when ||v_after|| < 0.7071 ||v_before||);
 * device-side decision, no host sync.  Final ||v|| goes to scal[out_slot];
 * if hsum is non-NULL the summed coefficients (k) are written there.
 * k == 0: only the norm. */
int bbjd_k_cgs2(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, void *v,
                bbjd_red *rd, int out_slot, void *hsum);

/* bbjd_k_cgs2, then v /= ||v|| when ||v|| > 1e-300 (scal[out_slot] keeps the
 * norm before the scaling). */
int bbjd_k_cgs2_unit(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, void *v,
                     bbjd_red *rd, int out_slot);

/* Test-space expansion: w = bbjd_k_cgs2_unit of (ar + i ai) a + (br + i bi) b
 * against W (a = A v, b = B v; w is written only as the result), which also
 * yields the projected pencil column colA[0..k] = [W w]^H a,
 * colB[0..k] = [W w]^H b (device arrays).  Returns 1, having done nothing,
 * when the shape takes the unfused path. */
int bbjd_k_cgs2_unit_pencil(hipStream_t s, int64_t n, int k, const void *W, int64_t ldw, void *w, const void *a,
                            const void *b, double are, double aim, double bre, double bim, void *colA, void *colB,
                            bbjd_red *rd, int out_slot);

/* v -= Q Q^H v twice (no norm test), k columns. */
int bbjd_k_project_out(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, void *v,
                       bbjd_red *rd);

/* out[j] = Q[:, j]^H x for j < k, and out2[j] = Q[:, j]^H x2 when x2 != NULL
 * (one pass over Q).  out/out2 are device arrays. */
int bbjd_k_gemv_c(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq,
                  const void *x, const void *x2, void *out, void *out2, bbjd_red *rd);

/* complex64 form of bbjd_k_gemv_c (FP32 accumulation) */
int bbjd_k_gemv_c_c(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq,
                    const void *x, const void *x2, void *out, void *out2, bbjd_red *rd);

/* y = Q[:, :k] c (c device, k entries) and, with slot >= 0, scal[slot] = ||y||. */
int bbjd_k_gemv_n(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq,
                  const void *c, void *y, bbjd_red *rd, int slot);

/* y = Q[:, :k] c for host coefficients c (k <= 128), passed in the kernel
 * arguments: no upload. */
int bbjd_k_gemv_n_host(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, const void *c, void *y);

/* ||x|| -> scal[slot] */
int bbjd_k_norm(hipStream_t s, int64_t n, const void *x, bbjd_red *rd, int slot);

/* mode 0: x *= 1/scal[slot]; mode 1: only when scal[slot] > 1e-300 */
int bbjd_k_scale_inv(hipStream_t s, int64_t n, void *x, const double *scal, int slot, int mode);

/* complex64 form of bbjd_k_scale_inv */
int bbjd_k_scale_inv_c(hipStream_t s, int64_t n, void *x, const double *scal, int slot, int mode);

/* bbjd_k_scale_inv(_c) that also writes the result to y (x unscaled when
 * mode 1 skips the scaling) */
int bbjd_k_scale_inv2(hipStream_t s, int64_t n, void *x, void *y, const double *scal, int slot, int mode);
int bbjd_k_scale_inv2_c(hipStream_t s, int64_t n, void *x, void *y, const double *scal, int slot, int mode);

/* x *= a (host scalar a = (are, aim)) */
int bbjd_k_scale(hipStream_t s, int64_t n, void *x, double are, double aim);

/* out = a x + b y (host scalars); out may alias x or y */
int bbjd_k_lincomb(hipStream_t s, int64_t n, double are, double aim, const void *x,
                   double bre, double bim, const void *y, void *out);

/* out = a x + b y with the coefficients in device memory:
 * co = {Re a, Im a, Re b, Im b} (doubles); same arithmetic as bbjd_k_lincomb.
 * _c: complex64 vectors. */
int bbjd_k_lincomb_dev(hipStream_t s, int64_t n, const double *co, const void *x, const void *y, void *out);
int bbjd_k_lincomb_dev_c(hipStream_t s, int64_t n, const double *co, const void *x, const void *y, void *out);

/* r = a x + b y and scal[slot] = ||r|| */
int bbjd_k_lincomb_norm(hipStream_t s, int64_t n, double are, double aim, const void *x,
                        double bre, double bim, const void *y, void *r, bbjd_red *rd, int slot);

/* Sleijpen-van der Vorst step: z -= s (u^H z)/(u^H s) when |u^H s| > 1e-300 */
int bbjd_k_svdv(hipStream_t s, int64_t n, const void *u, void *z, const void *sv, bbjd_red *rd);

/* X[:, :keep] = X[:, :k] Zr where Zr is device, row-major k x keep
 * (Zr[j keep + c] = Z[j, c]); in place, row-tiled. */
int bbjd_k_rotate(hipStream_t s, int64_t n, int k, int keep, void *X, int64_t ldx,
                  const void *Zr);

/* G (p x q, column-major, leading dimension ldg) = X^H Y for tall-skinny
 * X (n x p) and Y (n x q); deterministic (fixed-order chunk sums). */
int bbjd_k_gram(hipStream_t s, int64_t n, int p, const void *X, int64_t ldx, int q, const void *Y, int64_t ldy,
                void *G, int64_t ldg, bbjd_red *rd);

/* x -= Y[:, :k] (M c) in complex64 with FP32 accumulation: M is k x k,
 * column-major, c has k entries (device arrays) */
int bbjd_k_update_pre_c(hipStream_t s, int64_t n, int k, const void *Y, int64_t ldy, const void *M, const void *c,
                        void *x);

/* precision conversion */
int bbjd_k_z2c(hipStream_t s, int64_t n, const void *z, void *c);
int bbjd_k_c2z(hipStream_t s, int64_t n, const void *c, void *z);

/* ---- complex64 / FP32-accumulation kernels for the inner solve --------- */
int bbjd_k_cgs2_c(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, void *v,
                  bbjd_red *rd, int out_slot, void *hsum);
int bbjd_k_project_out_c(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq, void *v,
                         bbjd_red *rd);
int bbjd_k_gemv_n_c(hipStream_t s, int64_t n, int k, const void *Q, int64_t ldq,
                    const void *c, void *y, bbjd_red *rd, int slot);
int bbjd_k_norm_c(hipStream_t s, int64_t n, const void *x, bbjd_red *rd, int slot);
int bbjd_k_scale_c(hipStream_t s, int64_t n, void *x, double are, double aim);
int bbjd_k_lincomb_c(hipStream_t s, int64_t n, double are, double aim, const void *x,
                     double bre, double bim, const void *y, void *out);

/* Process-wide order between HIP graph captures and device-wide calls.
 * ROCm's allocation and release calls synchronise the whole device, which
 * invalidates a stream capture in progress on another thread: captures hold
 * the lock shared (bbjd_capture_*), set-up and release of workspaces and
 * operators exclusively (bbjd_devwide_*).  bbjd_devwide_lock takes it only
 * while a workspace with graphs exists (no capture can run otherwise) and
 * returns whether it did; pass that to bbjd_devwide_unlock. */
void bbjd_capture_lock(void);
void bbjd_capture_unlock(void);
int  bbjd_devwide_lock(void);
void bbjd_devwide_unlock(int held);

/* Launch recording for grouped execution of independent solves (launch.h,
 * kernels.hip).  Between bbjd_rec_begin and bbjd_rec_end, launches made by
 * the calling thread through the engine's launchers (and the copies below)
 * are appended to the job's list instead of being launched; bbjd_rec_flush
 * launches every job's list on s, merging launches of the same kernel from
 * different jobs into one grouped launch (each job's blocks see the grid
 * and arguments of its own launch).  *launches: launches made. */
void *bbjd_rec_create(void);
void  bbjd_rec_destroy(void *rec);
void  bbjd_rec_begin(void *rec, int job);
void  bbjd_rec_end(void);
int   bbjd_rec_flush(void *rec, hipStream_t s, long *launches);
/* the launch order of a flush for nj lists of lens[j] launch keys (keys of
 * list j at keys[sum of lens[0..j-1]], 0: a launch that is never merged):
 * groups receives each launch as its list count followed by the lists; returns
 * the number of launches, -1 when cap ints do not hold them (tests) */
int   bbjd_rec_plan(int nj, const int *lens, const int64_t *keys, int *groups, int cap);
/* launches recorded and launches made by flushes, since creation */
void  bbjd_rec_counts(const void *rec, long *recorded, long *launches);
/* hipMemcpyAsync / hipMemsetAsync, recorded in order while recording */
int   bbjd_copy_async(void *dst, const void *src, size_t bytes, int kind, hipStream_t s);
int   bbjd_memset_async(void *dst, int value, size_t bytes, hipStream_t s);
/* device copy of bytes (a multiple of 8) by a kernel (groupable) */
int   bbjd_k_copy(hipStream_t s, void *dst, const void *src, size_t bytes);

/* batcher job choice: 1 (default) the longest-waiting inner solve and those
 * nearest to it in n; 0 arrival order */
void  bbjd_set_batch_pick(int nearest_n);

/* bytes of reduction scratch for the given shape */
int64_t bbjd_red_part_entries(int64_t n, int maxk);

#ifdef __cplusplus
}
#endif
#endif
