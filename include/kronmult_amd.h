/*
 * kronmult_amd.h -- C ABI of the MI355X (gfx950) batched Kronecker mat-vec.
 *
 *     Y(:,k) += kron(A_1(k), ..., A_d(k)) * X(:,k),   k = 0 .. batchCount-1,
 *     d = 1..6.
 *
 * Conventions are the reference's (KRONMULT):
 *   - column-major storage; X and Y are tensors with the last index fastest,
 *     A_k(i,j) = A_k[i + j*ld] (kroncommon.hpp:72-152, kgemm_nt.hpp:81);
 *   - every call ACCUMULATES into Y; outputs of different entries may overlap
 *     (several pY[k] equal), in which case all contributions are summed
 *     (kgemm_nt.hpp:184-187);
 *   - all arrays, including the pointer arrays pX/pY/pW/Aarray, live in
 *     device memory (test_kronmult6_xbatched.cpp:104-133, 268-271);
 *   - the library allocates nothing on these calls; launches are asynchronous
 *     on `stream` (a hipStream_t, NULL = default stream), like the
 *     reference's null-stream launches (kronmult6_batched.cpp:12-14).
 *
 * Differences from the reference, all of them supersets:
 *   - every function returns a status instead of void, after validating its
 *     scalar arguments on the host;
 *   - X is never modified (the reference uses X as scratch in batched and
 *     pbatched, kronm_backward.hpp:128);
 *   - W / pW are only touched when the requested shape has no on-chip kernel
 *     (see km_query_kernel); they may be NULL otherwise;
 *   - options select the rounding order (bitwise reproduction of either
 *     reference build), a deterministic output-sorted reduction (plans), and
 *     the stage precision.
 */
#ifndef KRONMULT_AMD_H
#define KRONMULT_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum km_status
{
  KM_SUCCESS = 0,
  KM_ERROR_INVALID_VALUE = 1,          /* d, n, lda, batchCount, null pointer */
  KM_ERROR_INSUFFICIENT_WORKSPACE = 2, /* fallback kernel needs W / pW      */
  KM_ERROR_NOT_SUPPORTED = 3,          /* shape/option combination not built */
  KM_ERROR_HIP = 4,                    /* HIP runtime / launch failure       */
  KM_ERROR_PLAN_MISMATCH = 5           /* plan built for another call shape  */
} km_status;

/* Order in which the d mode contractions are applied (decides rounding). */
typedef enum km_variant
{
  KM_VARIANT_BACKWARD = 0,  /* modes d..1: Alg. 993 backward,
                               kronm_backward.hpp:92-139 (default)          */
  KM_VARIANT_FORWARD = 1,   /* modes 1..d: Alg. 993 forward / the recursive
                               default build, kronmultv_recur.hpp:90-145     */
  KM_VARIANT_ALG993 = 2     /* the reference's flop-based choice between the
                               two, kronm.hpp:83-107 (square => backward)   */
} km_variant;

/* Arithmetic of the stage contractions.  The two matrix-core values are
 * built for the two-pass shapes at n = 8 (d = 6, double data, fused
 * multiply-add calls); other calls with them return KM_ERROR_NOT_SUPPORTED.
 * They change the first pass (modes d..3 of every sub-slab); the last two
 * modes and the accumulation into Y stay FP64 vector arithmetic. */
typedef enum km_precision
{
  KM_PRECISION_FP64 = 0,      /* stages and accumulation in the data type    */
  KM_PRECISION_FP64_MFMA = 1, /* FP64 matrix instructions for the stages     */
  KM_PRECISION_F16X4 = 2      /* mixed: each operand split into four 10-bit
                                 FP16 slices under a power-of-two scale (per
                                 fiber of X, per row of A_k), FP16 matrix
                                 instructions with exact FP32 sums, the slice
                                 pairs recombined in FP64; about 40 bits kept
                                 per stage (error ~1e-12 relative to the
                                 largest element of the fiber and row)       */
} km_precision;

/* option flags */
#define KM_FLAG_EXACT 0x1u        /* separate rounded multiply and add in every
                                     dot product (the reference's CPU rounding);
                                     default uses fused multiply-add          */
#define KM_FLAG_FORCE_GLOBAL 0x2u /* always use the workspace-staged kernel   */
#define KM_FLAG_PERSISTENT 0x4u   /* grouped calls: one persistent launch over
                                     all sizes instead of one launch per size;
                                     outputs owned (as DETERMINISTIC)         */
#define KM_FLAG_DETERMINISTIC 0x8u /* grouped calls: every output owned by one
                                      workgroup, no atomics, bitwise
                                      reproducible run to run                */
#define KM_FLAG_PLAN_OUTPUTS 0x10u /* calls with an output-ordered plan (or a
                                      group plan in an owned mode) write to
                                      the outputs recorded when the plan was
                                      created and do not read pY; without it
                                      pY is checked on the device (see
                                      km_plan_create)                        */

typedef struct km_plan_s *km_plan;

typedef struct km_options
{
  unsigned flags;    /* KM_FLAG_* */
  int variant;       /* km_variant */
  int precision;     /* km_precision */
  km_plan plan;      /* optional deterministic reduction plan (km_plan_create) */
  int reserved[4];
} km_options;

/* Fill *opt with the defaults (backward order, FMA, FP64, no plan). */
void km_options_init(km_options *opt);

const char *km_status_string(km_status s);

/* Detail of the last KM_ERROR_HIP on the calling thread (failing HIP call and
 * its error string); "" when none occurred. */
const char *km_last_error(void);

/* ---------------------------------------------------------------------------
 * kronmultN_batched  (kronmult6_batched.cpp:4-18, kronmult6_batched.hpp:12-75)
 *   A: (n, n, d, batchCount) contiguous; X, Y, W: (n^d, batchCount)
 *   contiguous.  Y(:,k) += kron(A(:,:,1,k), ..., A(:,:,d,k)) X(:,k).
 * ------------------------------------------------------------------------ */
km_status km_kronmult_batched_f64(int d, int n, const double *A,
                                  const double *X, double *Y, double *W,
                                  int batchCount, void *stream,
                                  const km_options *opt);
km_status km_kronmult_batched_f32(int d, int n, const float *A, const float *X,
                                  float *Y, float *W, int batchCount,
                                  void *stream, const km_options *opt);

/* ---------------------------------------------------------------------------
 * kronmultN_pbatched  (kronmult6_pbatched.cpp:4-19, kronmult6_pbatched.hpp:12-67)
 *   A: (n, n, d, batchCount) contiguous; pX/pY/pW: device arrays of
 *   batchCount device pointers to vectors of n^d.  pY entries may repeat.
 * ------------------------------------------------------------------------ */
km_status km_kronmult_pbatched_f64(int d, int n, const double *A,
                                   double *const *pX, double *const *pY,
                                   double *const *pW, int batchCount,
                                   void *stream, const km_options *opt);
km_status km_kronmult_pbatched_f32(int d, int n, const float *A,
                                   float *const *pX, float *const *pY,
                                   float *const *pW, int batchCount,
                                   void *stream, const km_options *opt);

/* ---------------------------------------------------------------------------
 * kronmultN_xbatched  (kronmult6_xbatched.cpp:4-28, kronmult_xbatched.hpp:17-162)
 *   Aarray: device array of d*batchCount device pointers, A_k of entry b at
 *   Aarray[k + d*b], each n x n with leading dimension lda >= n (pointers may
 *   repeat: pooled A).  pX/pY: batchCount pointers (both may repeat).
 *   pW: workspace pairs as in the reference (pW[2*s], pW[2*s+1] for
 *   sub-batch slot s < subbatchCount; subbatchCount 0 means
 *   max(1, batchCount/2), kronmult6_xbatched.cpp:16-20).
 * ------------------------------------------------------------------------ */
km_status km_kronmult_xbatched_f64(int d, int n, const double *const *Aarray,
                                   int lda, double *const *pX,
                                   double *const *pY, double *const *pW,
                                   int batchCount, int subbatchCount,
                                   void *stream, const km_options *opt);
km_status km_kronmult_xbatched_f32(int d, int n, const float *const *Aarray,
                                   int lda, float *const *pX, float *const *pY,
                                   float *const *pW, int batchCount,
                                   int subbatchCount, void *stream,
                                   const km_options *opt);

/* ---------------------------------------------------------------------------
 * kronmultN_vbatched  (kronmult6_vbatched.hpp:10-34, kronmult_vbatched.hpp:17-228)
 *   Rectangular A_k (m[k] x n[k], leading dimension m[k]), shapes uniform
 *   over the batch.  Aarray: d*batchCount device pointers.  X of entry b has
 *   prod(n) elements, Y prod(m).  W: one workspace of Wcapacity_bytes, used
 *   only by the workspace-staged kernel (two slots of max stage size per
 *   concurrent entry, kronmult_vbatched.hpp:114-135).  The reference has no
 *   host wrapper for vbatched; this is that wrapper.
 * ------------------------------------------------------------------------ */
km_status km_kronmult_vbatched_f64(int d, const int *m, const int *n,
                                   const double *const *Aarray,
                                   double *const *pX, double *const *pY,
                                   double *W, size_t Wcapacity_bytes,
                                   int batchCount, void *stream,
                                   const km_options *opt);
km_status km_kronmult_vbatched_f32(int d, const int *m, const int *n,
                                   const float *const *Aarray, float *const *pX,
                                   float *const *pY, float *W,
                                   size_t Wcapacity_bytes, int batchCount,
                                   void *stream, const km_options *opt);

/* ---------------------------------------------------------------------------
 * Deterministic output-sorted reduction.  A plan records, for one set of
 * output pointers pY (device array of batchCount pointers), a stable order of
 * the entries grouped by output and a split of those groups over
 * workgroups, so that every output vector is owned by exactly one workgroup:
 * no atomics, bitwise reproducible results run to run.  No counterpart in
 * the reference (its overlapping outputs use atomics, kgemm_nt.hpp:184-187).
 * Creating a plan sorts the entries by output on the device (a stable radix
 * sort; batches above 4,194,304 entries are sorted on the host), keeps a
 * device copy of pY, allocates a few device arrays and synchronizes
 * `stream`; a call that passes the plan allocates nothing.
 * Binding: a plan is built for the pY contents it was created from.  Every
 * call with an output-ordered plan first compares the call's pY with that
 * copy on the device (one small kernel on `stream`); if any pointer differs,
 * the call accumulates into the caller's pY with atomics instead (correct,
 * not bitwise reproducible).  With KM_FLAG_PLAN_OUTPUTS the check is skipped
 * and the call writes to the outputs recorded in the plan (pY is not read).
 * A plan may be used from several streams at once.
 * elem_bytes is sizeof(double) or sizeof(float).
 * ------------------------------------------------------------------------ */
km_status km_plan_create(km_plan *plan, int d, int n, int elem_bytes,
                         const void *const *pY, int batchCount, void *stream);

/* Plan orders.  KM_ORDER_BY_OUTPUT is km_plan_create's (owned outputs,
 * deterministic).  KM_ORDER_BY_INPUT sorts entries by their X pointer so that
 * the entries reading the same input run back to back (the repeated reads
 * hit L2: pooled-input batches such as sparse-grid operator applies); the
 * outputs are then shared between workgroups and flushed with atomics, so
 * the result is not bitwise reproducible. */
typedef enum km_plan_order
{
  KM_ORDER_BY_OUTPUT = 0,
  KM_ORDER_BY_INPUT = 1
} km_plan_order;

km_status km_plan_create_ordered(km_plan *plan, int d, int n, int elem_bytes,
                                 const void *const *pX, const void *const *pY,
                                 int batchCount, int order_by, void *stream);
km_status km_plan_destroy(km_plan plan);

/* ---------------------------------------------------------------------------
 * Allocation-free deterministic pbatched.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Computes
 * what km_plan_create + km_kronmult_pbatched_f64 with that plan +
 * km_plan_destroy compute, with the same kernels and bitwise the same
 * results, but allocates nothing: the output-ordered entry order is built on
 * the device, on `stream`, in a static buffer of 8 MiB that each device the
 * library runs on holds until the process ends.
 *   - batchCount <= KM_NOALLOC_MAX_BATCH; larger batches return
 *     KM_ERROR_INVALID_VALUE (km_last_error says why);
 *   - opt->plan must be NULL (the call builds its own order);
 *   - every call on one device shares that device's buffer: calls that
 *     overlap on one device (two streams, two host threads) race and must be
 *     serialized by the caller;
 *   - pW as for km_kronmult_pbatched_f64 (d = 6 with n >= 7 needs it).
 * ------------------------------------------------------------------------ */
#define KM_NOALLOC_MAX_BATCH 262144
km_status km_kronmult_pbatched_noalloc_f64(int d, int n, const double *A,
                                           double *const *pX,
                                           double *const *pY,
                                           double *const *pW, int batchCount,
                                           void *stream,
                                           const km_options *opt);
/* number of distinct outputs and of workgroups in the plan */
km_status km_plan_info(km_plan plan, int *num_outputs, int *num_groups);

/* ---------------------------------------------------------------------------
 * Grouped variable-size batch: entries with their own square size n_b
 * (adaptive sparse grids), all with the same d, in ONE persistent launch.
 * Extends the reference's uniform-shape vbatched (kronmult_vbatched.hpp:
 * 19-24, where every entry shares m_k, n_k) to per-entry sizes; there is no
 * counterpart in the reference.
 *   Aarray: d*batchCount device pointers, A_k of entry b at Aarray[k + d*b],
 *           n_b x n_b with leading dimension n_b;
 *   pX/pY:  batchCount device pointers to vectors of n_b^d elements
 *           (outputs may repeat; entries sharing an output must have the
 *           same n_b).
 * The group plan bins the entries by size and sorts each bin by output.  By
 * default a call issues one on-chip launch per size bin (back to back on
 * `stream`, capturable in a HIP graph), each bin split evenly over the
 * resident workgroups, outputs shared between workgroups accumulated with
 * atomics.  KM_FLAG_DETERMINISTIC splits each bin at output boundaries
 * instead, so every output is owned by one workgroup (no atomics, bitwise
 * reproducible); its bins run concurrently on up to four streams the plan
 * owns, forked from `stream` and joined back into it with events (also
 * capturable; without KM_FLAG_EXACT the small sizes go together into one
 * launch of the one-launch kernel below); calls from several host threads
 * on one plan take turns through the fork (a per-plan lock).  KM_FLAG_PERSISTENT runs all bins in
 * one launch with owned outputs (each workgroup takes whole output runs;
 * bitwise reproducible; the fastest owned mode on mixed sizes, within about
 * 1.05x of the default on a d = 4 mix of n = 2..8).  An output run of large entries too long for one workgroup is
 * cut into parts whose sums the plan keeps in its own device buffer and a
 * second short launch adds to the output in part order, so such a call is
 * two launches, and persistent calls of one plan must not overlap (on
 * several streams) while the plan has such runs (num_groups counts the
 * first launch's workgroups).  Creating the plan reads the host array n_host
 * and the device pointer array pY once.  The owned modes are bound to those
 * pY contents as output-ordered plans are (km_plan_create): a call whose pY
 * differs accumulates with atomics, and KM_FLAG_PLAN_OUTPUTS writes to the
 * recorded outputs.
 * Supported sizes: n_b <= 8 for d <= 4, n_b <= 6 for d = 5, n_b <= 4 for
 * d = 6 (KM_ERROR_NOT_SUPPORTED otherwise).
 * ------------------------------------------------------------------------ */
typedef struct km_group_plan_s *km_group_plan;

km_status km_group_plan_create(km_group_plan *plan, int d, int batchCount,
                               const int *n_host, const void *const *pY,
                               int elem_bytes, void *stream);
km_status km_group_plan_destroy(km_group_plan plan);
km_status km_group_plan_info(km_group_plan plan, int *num_outputs,
                             int *num_groups, int *max_n);
km_status km_kronmult_grouped_f64(km_group_plan plan,
                                  const double *const *Aarray,
                                  double *const *pX, double *const *pY,
                                  void *stream, const km_options *opt);
km_status km_kronmult_grouped_f32(km_group_plan plan,
                                  const float *const *Aarray, float *const *pX,
                                  float *const *pY, void *stream,
                                  const km_options *opt);

/* ---------------------------------------------------------------------------
 * Kernel-variant tuning.  Several kernels can serve one square (d, n) call
 * with fused multiply-adds; the library keeps a table from (d, n, element
 * size, owned outputs (an output-ordered plan) or not, batch-size bin
 * floor(log2 batchCount)) to a variant, and uses its built-in choice where
 * the table has no entry.  No counterpart in the reference (its launch
 * geometry is fixed, kronmult6_batched.cpp:8-14).
 *   km_tune_shape  times every candidate variant on synthetic device data of
 *                  the given shape (allocates, fills and frees device memory,
 *                  synchronizes `stream`; outputs shared `overlap`:1) and
 *                  records the fastest in the table;
 *   km_tune_set / km_tune_get  write / read one entry (KM_KERNEL_DEFAULT
 *                  removes it);
 *   km_tune_save / km_tune_load  the table as JSON, keyed by GPU architecture
 *                  (a file tuned on another architecture is refused with
 *                  KM_ERROR_NOT_SUPPORTED).  The file named by the environment
 *                  variable KRONMULT_AMD_TUNE_FILE is loaded on first use.
 * EXACT calls keep the built-in choice (it fixes the rounding order).
 * ------------------------------------------------------------------------ */
typedef enum km_kernel_variant
{
  KM_KERNEL_DEFAULT = 0,  /* built-in choice                                 */
  KM_KERNEL_SLAB = 1,     /* slab-streamed, one workgroup per entry          */
  KM_KERNEL_HALF1 = 2,    /* slab-streamed, two workgroups per entry split by
                             the last mode's rows, one slab per step        */
  KM_KERNEL_HALF2 = 3,    /* the same, two slabs per step                    */
  KM_KERNEL_REG = 4,      /* register-blocked mode groups                    */
  KM_KERNEL_LDS = 5,      /* whole tensor in LDS                             */
  KM_KERNEL_LDS_WIDE = 6, /* whole tensor in LDS, wider workgroups           */
  KM_KERNEL_TWOPASS = 7,  /* n^(d-1) above LDS (d = 6, n = 7, 8): two
                             kernels, W through HBM                         */
  KM_KERNEL_XCD = 8       /* the same in one cooperative launch, W in each
                             XCD's L2 (gfx950); half the HBM bytes, but
                             slower where measured, so never the built-in
                             choice                                         */
} km_kernel_variant;

km_status km_tune_shape(int d, int n, int elem_bytes, int batchCount,
                        int overlap, int owned, void *stream, int *variant);
km_status km_tune_set(int d, int n, int elem_bytes, int batchCount, int owned,
                      int variant);
km_status km_tune_get(int d, int n, int elem_bytes, int batchCount, int owned,
                      int *variant);
km_status km_tune_clear(void);
km_status km_tune_save(const char *path);
km_status km_tune_load(const char *path);

/* ---------------------------------------------------------------------------
 * Introspection: which kernel family a square (d, n) call would use.
 *   *family: 0 = LDS-resident, 1 = slab-streamed, 2 = workspace-staged
 *            (generic), 3 = two-pass (sub-slabs in LDS, then the two
 *            leading modes in registers; uses W / pW), 4 = register-blocked
 *            (mode groups in registers, LDS transposes between them),
 *            7 = the two-pass family fused into one cooperative launch on
 *            gfx950 (each XCD runs its share of the entries with W in its
 *            L2; FMA calls whose entries each bring a workspace), used where
 *            the variant table selects KM_KERNEL_XCD.
 *   *threads: workgroup size.
 * ------------------------------------------------------------------------ */
km_status km_query_kernel(int d, int n, int elem_bytes, const km_options *opt,
                          int *family, int *threads);
/* The same for a call of batchCount entries, with (owned = 1) or without an
 * output-ordered plan: the tuned variant for that batch bin applies.
 * *family 6 = split pair (two workgroups per entry). */
km_status km_query_kernel_for(int d, int n, int elem_bytes, int batchCount,
                              int owned, const km_options *opt, int *family,
                              int *threads);

/* Launches of the fused two-pass kernel (family 7) on the current device
 * whose bounded inter-workgroup waits gave up, since the library was loaded
 * (synchronous).  Such a launch exits without completing its outputs; it
 * cannot happen while every workgroup of the cooperative launch is resident,
 * which the launch guarantees.  No counterpart in the reference (whose
 * two-pass chain is separate kernels, kronm_backward.hpp:92-139); a check
 * for callers and tests of the fused path. */
km_status km_cooperative_faults(unsigned long long *count);

/* Reference flop model (kron_flops, kroncommon.hpp:154-185) for one entry. */
double km_kron_flops(int d, const int *m, const int *n);

/* ---------------------------------------------------------------------------
 * Pooled xbatched tables from index lists, built on the device: the step
 * before km_kronmult_xbatched_* in a sparse-grid operator apply, where each
 * entry b names one matrix per dimension from a pool, one input vector and
 * one output vector.  No counterpart in the reference (its callers build the
 * tables); the A table has the layout kronmultN_xbatched reads,
 * Aarray[k + d*b] (kronmult_xbatched.hpp:67-69).
 *   Apool, a_stride, a_count: HOST arrays of d entries: device base of pool
 *     k, elements between consecutive matrices of pool k, matrices in pool k;
 *   aidx: device int[d*batch], matrix of dimension k for entry b at [k + d*b];
 *   X/x_stride/x_count, xidx: input pool (device), element stride between
 *     vectors, vectors in the pool, device int[batch] index per entry;
 *   Y/y_stride/y_count, yidx: the same for the outputs (indices may repeat);
 *   Aarray (d*batch), pX, pY (batch): device outputs;
 *   bad: optional device int, incremented once per index outside its pool
 *     (that pointer is written as NULL).
 * Asynchronous on `stream`; allocates nothing.
 * ------------------------------------------------------------------------ */
km_status km_xbatched_tables_f64(int d, int batchCount, const double *const *Apool,
                                 const long long *a_stride, const int *a_count,
                                 const int *aidx, double *X, long long x_stride,
                                 int x_count, const int *xidx, double *Y,
                                 long long y_stride, int y_count,
                                 const int *yidx, const double **Aarray,
                                 double **pX, double **pY, int *bad,
                                 void *stream);
km_status km_xbatched_tables_f32(int d, int batchCount, const float *const *Apool,
                                 const long long *a_stride, const int *a_count,
                                 const int *aidx, float *X, long long x_stride,
                                 int x_count, const int *xidx, float *Y,
                                 long long y_stride, int y_count,
                                 const int *yidx, const float **Aarray,
                                 float **pX, float **pY, int *bad,
                                 void *stream);

/* ---------------------------------------------------------------------------
 * Batched slender GEMM:  C_b = alpha*op(A_b)*op(B_b) + beta*C_b,  b < batchCount.
 * Replaces the reference's kgemm_nn_batched / kgemm_nt_batched /
 * kgemm_tt_batched (host wrappers kgemm_nn_batched.cpp:5-25,
 * kgemm_nt_batched.cpp:5-25, kgemm_tt_batched.cpp:5-25; templates
 * kgemm_nn_batched.hpp:7-63 and copies).
 *   nn: A mm x kk, B kk x nn;  nt: A mm x kk, B nn x kk (C = A*B^T);
 *   tt: A kk x mm, B nn x kk (C = A^T*B^T).  C is mm x nn.
 *   Aarray/Barray/Carray and the leading-dimension arrays are device arrays
 *   of batchCount entries (as in the reference); column-major storage.
 *   beta == 1: C += alpha*(...) with atomic adds, so entries may share C;
 *   beta == 0: C = alpha*(...), C not read;
 *   otherwise C = beta*C + alpha*(...) (entries must not share C).
 * Each element is a dot product over k ascending in double (float data is
 * accumulated in double, kgemm_nn.hpp:214-221).  KM_FLAG_EXACT uses separate
 * rounded multiplies and adds (the reference CPU build's rounding); the
 * default fuses them.  Asynchronous on `stream`; allocates nothing.
 * ------------------------------------------------------------------------ */
km_status km_kgemm_nn_batched_f64(int mm, int nn, int kk, double alpha,
                                  const double *const *Aarray,
                                  const int *ldAarray,
                                  const double *const *Barray,
                                  const int *ldBarray, double beta,
                                  double *const *Carray, const int *ldCarray,
                                  int batchCount, void *stream,
                                  const km_options *opt);
km_status km_kgemm_nt_batched_f64(int mm, int nn, int kk, double alpha,
                                  const double *const *Aarray,
                                  const int *ldAarray,
                                  const double *const *Barray,
                                  const int *ldBarray, double beta,
                                  double *const *Carray, const int *ldCarray,
                                  int batchCount, void *stream,
                                  const km_options *opt);
km_status km_kgemm_tt_batched_f64(int mm, int nn, int kk, double alpha,
                                  const double *const *Aarray,
                                  const int *ldAarray,
                                  const double *const *Barray,
                                  const int *ldBarray, double beta,
                                  double *const *Carray, const int *ldCarray,
                                  int batchCount, void *stream,
                                  const km_options *opt);
km_status km_kgemm_nn_batched_f32(int mm, int nn, int kk, float alpha,
                                  const float *const *Aarray,
                                  const int *ldAarray,
                                  const float *const *Barray,
                                  const int *ldBarray, float beta,
                                  float *const *Carray, const int *ldCarray,
                                  int batchCount, void *stream,
                                  const km_options *opt);
km_status km_kgemm_nt_batched_f32(int mm, int nn, int kk, float alpha,
                                  const float *const *Aarray,
                                  const int *ldAarray,
                                  const float *const *Barray,
                                  const int *ldBarray, float beta,
                                  float *const *Carray, const int *ldCarray,
                                  int batchCount, void *stream,
                                  const km_options *opt);
km_status km_kgemm_tt_batched_f32(int mm, int nn, int kk, float alpha,
                                  const float *const *Aarray,
                                  const int *ldAarray,
                                  const float *const *Barray,
                                  const int *ldBarray, float beta,
                                  float *const *Carray, const int *ldCarray,
                                  int batchCount, void *stream,
                                  const km_options *opt);

#ifdef __cplusplus
}
#endif

#endif /* KRONMULT_AMD_H */
