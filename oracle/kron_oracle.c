/*
 * kron_oracle.c -- CPU restatement of KRONMULT's batched Kronecker
 * matrix-vector product  Y(:,k) += kron(A1(k),...,Ad(k)) * X(:,k).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker; the product path (kronmult_amd) never
 * links or calls it.
 *
 * It restates, loop for loop, the reference's CPU (OpenMP) build:
 *   - slender GEMM element loops: kgemm_nt.hpp:87-196, kgemm_tt.hpp:86-194,
 *     kgemm_nn.hpp:91-197.  Every element is a dot product over k ascending,
 *     started from 0, one multiply and one add per term (this file is built
 *     with -ffp-contract=off so no FMA contraction changes the rounding);
 *   - recursive variant (the reference's default build):
 *     kronmultv_recur.hpp:16-213;
 *   - Algorithm 993 backward / forward variants and the choice between them:
 *     kronm_backward.hpp:15-161, kronm_forward.hpp:15-167, kronm.hpp:18-108;
 *   - the final-stage Y update as an OpenMP atomic add: kroncommon.hpp:35-59;
 *   - batch drivers with an entry-level "omp parallel for":
 *     kronmult6_batched.hpp:12-75, kronmult6_pbatched.hpp:12-67,
 *     kronmult_xbatched.hpp:17-162 (X copy into workspace pairs, sub-batches),
 *     kronmult_vbatched.hpp:17-228 (one workspace buffer, capacity-driven
 *     sub-batches);
 *   - the brute-force Kronecker sum of the reference tests:
 *     test_kronmult6_xbatched.cpp:397-500;
 *   - the flop model kron_flops: kroncommon.hpp:154-185;
 *   - the standalone batched slender GEMMs kgemm_{nn,nt,tt}_batched
 *     (kgemm_nn_batched.hpp:7-63 and copies), with their alpha/beta rules.
 *
 * Float entry points follow the reference's float templates: products and
 * sums in double, every stage result rounded to float when stored
 * (kgemm_nt.hpp:213-221, kgemm_tt.hpp:211-219, kgemm_nn.hpp:214-221).
 *
 * Every top-level call uses nvec = 1, as every reference batch driver does
 * (kronmult6_pbatched.hpp:64), so the nvec > 1 transpose branches of
 * kronm_forward/backward are not restated.
 *
 * Indexing is 0-based here; the reference's Fortran helpers (indx2f..indx6f,
 * kroncommon.hpp:72-152) describe the same column-major layout.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum
{
  ORC_RECURSIVE = 0, /* reference default build */
  ORC_BACKWARD  = 1, /* Alg. 993 backward, forced */
  ORC_FORWARD   = 2, /* Alg. 993 forward, forced */
  ORC_ALG993    = 3  /* Alg. 993 with the reference's flop-based choice */
};

static int iprod(int nd, const int *v)
{
  int p = 1;
  for (int i = 0; i < nd; i++)
    p *= v[i];
  return p;
}

/* kronm.hpp:37-80 -- flop estimates of the two Alg. 993 variants */
static double flops_forward(int nd, const int *m, const int *n)
{
  int xs = iprod(nd, n);
  double tot = 0;
  for (int i = 0; i < nd; i++)
  {
    int mm = m[i], kk = n[i], nn = xs / kk;
    tot += (2.0 * mm) * nn * kk;
    xs = mm * nn;
  }
  return tot;
}

static double flops_backward(int nd, const int *m, const int *n)
{
  int xs = iprod(nd, n);
  double tot = 0;
  for (int i = nd - 1; i >= 0; i--)
  {
    int nn = m[i], kk = n[i], mm = xs / kk;
    tot += (2.0 * mm) * nn * kk;
    xs = mm * nn;
  }
  return tot;
}

/* kronm.hpp:89: forward only when strictly cheaper, so square => backward */
int oracle_alg993_uses_forward(int nd, const int *m, const int *n)
{
  return flops_forward(nd, m, n) < flops_backward(nd, m, n);
}

/* kroncommon.hpp:154-185 */
double oracle_kron_flops(int nd, const int *m, const int *n)
{
  if (nd == 1)
    return 2.0 * m[0] * n[0];
  double f1 = 2.0 * (double)iprod(nd - 1, n + 1) * n[0] * m[0];
  return f1 + oracle_kron_flops(nd - 1, m + 1, n + 1) * m[0];
}

/* -------------------------------------------------------------------------
 * Type-generic part: T is the storage type, ACC the dot-product type.
 * ---------------------------------------------------------------------- */
#define DEFINE_ORACLE(T, ACC, SUF)                                             \
                                                                               \
  static inline void atomic_add_##SUF(T *p, T v)                               \
  {                                                                            \
    _Pragma("omp atomic update") * p += v;                                     \
  }                                                                            \
                                                                               \
  /* kgemm_nt.hpp:87-196: C(i,j) = sum_k A(i,k)*B(j,k) */                      \
  static void gemm_nt_##SUF(int mm, int nn, int kk, const T *A, int ldA,       \
                            const T *B, int ldB, int acc, T *C, int ldC)       \
  {                                                                            \
    for (int j = 0; j < nn; j++)                                               \
      for (int i = 0; i < mm; i++)                                             \
      {                                                                        \
        ACC c = 0;                                                             \
        const T *ap = A + i;                                                   \
        const T *bp = B + j;                                                   \
        for (int k = 0; k < kk; k++)                                           \
        {                                                                      \
          ACC a = *ap, b = *bp;                                                \
          c += a * b;                                                          \
          ap += ldA;                                                           \
          bp += ldB;                                                           \
        }                                                                      \
        if (acc)                                                               \
          atomic_add_##SUF(&C[i + (size_t)j * ldC], (T)c);                     \
        else                                                                   \
          C[i + (size_t)j * ldC] = (T)c;                                       \
      }                                                                        \
  }                                                                            \
                                                                               \
  /* kgemm_tt.hpp:86-194: C(i,j) = sum_k A(k,i)*B(j,k) */                      \
  static void gemm_tt_##SUF(int mm, int nn, int kk, const T *A, int ldA,       \
                            const T *B, int ldB, int acc, T *C, int ldC)       \
  {                                                                            \
    for (int j = 0; j < nn; j++)                                               \
      for (int i = 0; i < mm; i++)                                             \
      {                                                                        \
        ACC c = 0;                                                             \
        const T *ap = A + (size_t)i * ldA;                                     \
        const T *bp = B + j;                                                   \
        for (int k = 0; k < kk; k++)                                           \
        {                                                                      \
          ACC a = *ap, b = *bp;                                                \
          c += a * b;                                                          \
          ap += 1;                                                             \
          bp += ldB;                                                           \
        }                                                                      \
        if (acc)                                                               \
          atomic_add_##SUF(&C[i + (size_t)j * ldC], (T)c);                     \
        else                                                                   \
          C[i + (size_t)j * ldC] = (T)c;                                       \
      }                                                                        \
  }                                                                            \
                                                                               \
  /* kgemm_nn.hpp:91-197: C(i,j) = sum_k A(i,k)*B(k,j) */                      \
  static void gemm_nn_##SUF(int mm, int nn, int kk, const T *A, int ldA,       \
                            const T *B, int ldB, int acc, T *C, int ldC)       \
  {                                                                            \
    for (int i = 0; i < mm; i++)                                               \
      for (int j = 0; j < nn; j++)                                             \
      {                                                                        \
        ACC c = 0;                                                             \
        const T *ap = A + i;                                                   \
        const T *bp = B + (size_t)j * ldB;                                     \
        for (int k = 0; k < kk; k++)                                           \
        {                                                                      \
          ACC a = *ap, b = *bp;                                                \
          c += a * b;                                                          \
          ap += ldA;                                                           \
          bp += 1;                                                             \
        }                                                                      \
        if (acc)                                                               \
          atomic_add_##SUF(&C[i + (size_t)j * ldC], (T)c);                     \
        else                                                                   \
          C[i + (size_t)j * ldC] = (T)c;                                       \
      }                                                                        \
  }                                                                            \
                                                                               \
  /* kronmultv_recur.hpp:16-146 (level) and :148-213 (base, Y += Ad*X) */      \
  static void recur_##SUF(int nd, const int *m, const int *n,                  \
                          const T *const *A, const int *ld, int nvec, T *X,    \
                          T *Y, T *W)                                          \
  {                                                                            \
    if (nd == 1)                                                               \
    {                                                                          \
      gemm_nn_##SUF(m[0], nvec, n[0], A[0], ld[0], X, n[0], 1, Y, m[0]);       \
      return;                                                                  \
    }                                                                          \
    int n26 = iprod(nd - 1, n + 1);                                            \
    size_t ldX = (size_t)n26 * n[0], ldW = (size_t)n26 * m[0];                 \
    for (int v = 0; v < nvec; v++)                                             \
      gemm_nt_##SUF(n26, m[0], n[0], X + v * ldX, n26, A[0], ld[0], 0,         \
                    W + v * ldW, n26);                                         \
    recur_##SUF(nd - 1, m + 1, n + 1, A + 1, ld + 1, nvec * m[0], W, Y, X);    \
  }                                                                            \
                                                                               \
  /* kronm_backward.hpp:77-139 with nvec = 1 */                                \
  static void backward_##SUF(int nd, const int *m, const int *n,               \
                             const T *const *A, const int *ld, T *X, T *Y,     \
                             T *W)                                             \
  {                                                                            \
    T *xin = X, *yout = W;                                                     \
    int xin_size = iprod(nd, n);                                               \
    for (int i = nd - 1; i >= 0; --i)                                          \
    {                                                                          \
      int fin = (i == 0);                                                      \
      int nn = m[i], kk = n[i], mm = xin_size / kk;                            \
      gemm_tt_##SUF(mm, nn, kk, xin, kk, A[i], ld[i], fin, fin ? Y : yout,     \
                    mm);                                                       \
      T *t = xin;                                                              \
      xin = yout;                                                              \
      yout = t;                                                                \
      xin_size = mm * nn;                                                      \
    }                                                                          \
  }                                                                            \
                                                                               \
  /* kronm_forward.hpp:81-164 with nvec = 1 */                                 \
  static void forward_##SUF(int nd, const int *m, const int *n,                \
                            const T *const *A, const int *ld, T *X, T *Y,      \
                            T *W)                                              \
  {                                                                            \
    T *xin = X, *yout = W;                                                     \
    int xin_size = iprod(nd, n);                                               \
    for (int i = 0; i < nd; i++)                                               \
    {                                                                          \
      int fin = (i == nd - 1);                                                 \
      int mm = m[i], kk = n[i], nn = xin_size / kk;                            \
      gemm_nt_##SUF(mm, nn, kk, A[i], ld[i], xin, nn, fin, fin ? Y : yout,     \
                    mm);                                                       \
      T *t = xin;                                                              \
      xin = yout;                                                              \
      yout = t;                                                                \
      xin_size = mm * nn;                                                      \
    }                                                                          \
  }                                                                            \
                                                                               \
  /* kronmultv.hpp:12-46 -> one entry, nvec = 1.  X is clobbered. */           \
  void oracle_entry_##SUF(int nd, const int *m, const int *n,                  \
                          const T *const *A, const int *ld, int variant,       \
                          T *X, T *Y, T *W)                                    \
  {                                                                            \
    if (nd == 1)                                                               \
    { /* kronmult1.hpp:38-51, kronmultv1.hpp: always Y += A1*X */              \
      gemm_nn_##SUF(m[0], 1, n[0], A[0], ld[0], X, n[0], 1, Y, m[0]);          \
      return;                                                                  \
    }                                                                          \
    switch (variant)                                                           \
    {                                                                          \
    case ORC_RECURSIVE:                                                        \
      recur_##SUF(nd, m, n, A, ld, 1, X, Y, W);                                \
      break;                                                                   \
    case ORC_BACKWARD:                                                         \
      backward_##SUF(nd, m, n, A, ld, X, Y, W);                                \
      break;                                                                   \
    case ORC_FORWARD:                                                          \
      forward_##SUF(nd, m, n, A, ld, X, Y, W);                                 \
      break;                                                                   \
    default:                                                                   \
      if (oracle_alg993_uses_forward(nd, m, n))                                \
        forward_##SUF(nd, m, n, A, ld, X, Y, W);                               \
      else                                                                     \
        backward_##SUF(nd, m, n, A, ld, X, Y, W);                              \
    }                                                                          \
  }                                                                            \
                                                                               \
  static void square_dims_##SUF(int nd, int n, int lda, int *m, int *nn,      \
                               int *ld)                                        \
  {                                                                            \
    for (int k = 0; k < nd; k++)                                               \
    {                                                                          \
      m[k] = n;                                                                \
      nn[k] = n;                                                               \
      ld[k] = lda;                                                             \
    }                                                                          \
  }                                                                            \
                                                                               \
  /* kronmult6_batched.hpp:12-75: A (n,n,d,batch), X/Y/W (n^d,batch) */       \
  void oracle_batched_##SUF(int nd, int n, const T *A, T *X, T *Y, T *W,       \
                            int batch, int variant)                            \
  {                                                                            \
    int m[6], nn[6], ld[6];                                                    \
    square_dims_##SUF(nd, n, n, m, nn, ld);                                          \
    size_t xs = (size_t)iprod(nd, nn);                                         \
    _Pragma("omp parallel for schedule(static)") for (int b = 0; b < batch;    \
                                                      b++)                     \
    {                                                                          \
      const T *Ap[6];                                                          \
      for (int k = 0; k < nd; k++)                                             \
        Ap[k] = A + ((size_t)k + (size_t)nd * b) * n * n;                      \
      oracle_entry_##SUF(nd, m, nn, Ap, ld, variant, X + b * xs, Y + b * xs,   \
                         W + b * xs);                                          \
    }                                                                          \
  }                                                                            \
                                                                               \
  /* kronmult6_pbatched.hpp:12-67: A (n,n,d,batch), pointer arrays X/Y/W */    \
  void oracle_pbatched_##SUF(int nd, int n, const T *A, T **pX, T **pY,        \
                             T **pW, int batch, int variant)                   \
  {                                                                            \
    int m[6], nn[6], ld[6];                                                    \
    square_dims_##SUF(nd, n, n, m, nn, ld);                                          \
    _Pragma("omp parallel for schedule(static)") for (int b = 0; b < batch;    \
                                                      b++)                     \
    {                                                                          \
      const T *Ap[6];                                                          \
      for (int k = 0; k < nd; k++)                                             \
        Ap[k] = A + ((size_t)k + (size_t)nd * b) * n * n;                      \
      oracle_entry_##SUF(nd, m, nn, Ap, ld, variant, pX[b], pY[b], pW[b]);     \
    }                                                                          \
  }                                                                            \
                                                                               \
  /* kronmult_xbatched.hpp:17-162: A pointer table (d,batch) + lda; X is     \
   * copied to pW[2*iz] and pW[2*iz+1] is the stage workspace of sub-batch    \
   * slot iz; with batch == 1 X itself is the workspace and pW[0] is W. */     \
  void oracle_xbatched_##SUF(int nd, int n, const T *const *Aptr, int lda,     \
                             T **pX, T **pY, T **pW, int batch, int subbatch,  \
                             int variant)                                      \
  {                                                                            \
    int m[6], nn[6], ld[6];                                                    \
    square_dims_##SUF(nd, n, lda, m, nn, ld);                                        \
    size_t xs = (size_t)iprod(nd, nn);                                         \
    if (subbatch == 0)                                                         \
      subbatch = (batch / 2 > 1) ? batch / 2 : 1;                              \
    for (int b0 = 0; b0 < batch; b0 += subbatch)                               \
    {                                                                          \
      int cnt = (batch - b0 < subbatch) ? batch - b0 : subbatch;               \
      _Pragma("omp parallel for schedule(static)") for (int iz = 0; iz < cnt;  \
                                                        iz++)                  \
      {                                                                        \
        int b = b0 + iz;                                                       \
        T *xp = (batch > 1) ? pW[2 * iz] : pX[b];                              \
        T *wp = (batch > 1) ? pW[2 * iz + 1] : pW[b];                          \
        if (xp != pX[b])                                                       \
          memcpy(xp, pX[b], xs * sizeof(T));                                   \
        const T *Ap[6];                                                        \
        for (int k = 0; k < nd; k++)                                           \
          Ap[k] = Aptr[k + (size_t)nd * b];                                    \
        oracle_entry_##SUF(nd, m, nn, Ap, ld, variant, xp, pY[b], wp);         \
      }                                                                        \
    }                                                                          \
  }                                                                            \
                                                                               \
  /* kronmult_vbatched.hpp:17-228: uniform rectangular shapes, ld_k = m_k,    \
   * one workspace W of Wcap bytes carved into (Xcopy, W) slot pairs of       \
   * sizeXW each.  Returns the sub-batch count used, or -1 when Wcap cannot    \
   * hold even one slot pair (the reference forces 1 and overruns W there,    \
   * kronmult_vbatched.hpp:131-135; the oracle refuses instead). */            \
  int oracle_vbatched_##SUF(int nd, const int *m, const int *n,                \
                            const T *const *Aptr, T **pX, T **pY, T *W,        \
                            size_t Wcap, int batch, int variant)               \
  {                                                                            \
    int ld[6];                                                                 \
    for (int k = 0; k < nd; k++)                                               \
      ld[k] = m[k];                                                            \
    size_t sx = (size_t)iprod(nd, n), sw = 0;                                  \
    for (int L = 1; L < nd; L++)                                               \
    { /* forward-order stage sizes (the reference's sizeW) and, so that the    \
         backward order can never overrun a slot, backward-order sizes too */  \
      size_t s = (size_t)iprod(L, m) * iprod(nd - L, n + L);                   \
      size_t sb = (size_t)iprod(nd - L, n) * iprod(L, m + nd - L);             \
      if (s > sw)                                                              \
        sw = s;                                                                \
      if (sb > sw)                                                             \
        sw = sb;                                                               \
    }                                                                          \
    size_t sxw = sx > sw ? sx : sw;                                            \
    size_t cap = Wcap / (2 * sxw * sizeof(T));                                 \
    if (cap < 1)                                                               \
      return -1;                                                               \
    int sub = (size_t)batch < cap ? batch : (int)cap;                          \
    for (int b0 = 0; b0 < batch; b0 += sub)                                    \
    {                                                                          \
      int cnt = (batch - b0 < sub) ? batch - b0 : sub;                         \
      _Pragma("omp parallel for schedule(static)") for (int iz = 0; iz < cnt;  \
                                                        iz++)                  \
      {                                                                        \
        int b = b0 + iz;                                                       \
        T *xp = W + (size_t)iz * 2 * sxw;                                      \
        T *wp = xp + sxw;                                                      \
        memcpy(xp, pX[b], sx * sizeof(T));                                     \
        const T *Ap[6];                                                        \
        for (int k = 0; k < nd; k++)                                           \
          Ap[k] = Aptr[k + (size_t)nd * b];                                    \
        oracle_entry_##SUF(nd, m, n, Ap, ld, variant, xp, pY[b], wp);          \
      }                                                                        \
    }                                                                          \
    return sub;                                                                \
  }                                                                            \
                                                                               \
  /* test_kronmult6_xbatched.cpp:447-499: explicit Kronecker sum, O(size^2) */ \
  void oracle_brute_##SUF(int nd, const int *m, const int *n,                  \
                          const T *const *A, const int *ld, const T *X, T *Y)  \
  {                                                                            \
    int ys = iprod(nd, m), xs = iprod(nd, n);                                  \
    _Pragma("omp parallel for schedule(static)") for (int ic = 0; ic < ys;     \
                                                      ic++)                    \
    {                                                                          \
      int iv[6], r = ic;                                                       \
      for (int k = nd - 1; k >= 0; k--)                                        \
      {                                                                        \
        iv[k] = r % m[k];                                                      \
        r /= m[k];                                                             \
      }                                                                        \
      double y = 0;                                                            \
      for (int jc = 0; jc < xs; jc++)                                          \
      {                                                                        \
        int jv[6], s = jc;                                                     \
        for (int k = nd - 1; k >= 0; k--)                                      \
        {                                                                      \
          jv[k] = s % n[k];                                                    \
          s /= n[k];                                                           \
        }                                                                      \
        double c = 1;                                                          \
        for (int k = 0; k < nd; k++)                                           \
          c *= (double)A[k][iv[k] + (size_t)jv[k] * ld[k]];                    \
        y += c * (double)X[jc];                                                \
      }                                                                        \
      Y[ic] = (T)((double)Y[ic] + y);                                          \
    }                                                                          \
  }

DEFINE_ORACLE(double, double, f64)
DEFINE_ORACLE(float, double, f32)

/*
 * Batched slender GEMM, the reference's kgemm_{nn,nt,tt}_batched
 * (kgemm_nn_batched.hpp:7-63, kgemm_nt_batched.hpp:7-62,
 * kgemm_tt_batched.hpp:7-62; host wrappers kgemm_nn_batched.cpp:5-25 and the
 * nt/tt copies).  Entries run in batch order (the CPU build's loop,
 * kgemm_nn_batched.hpp:53-62); each element of entry b is
 *   c = sum_k op(A)(i,k) * op(B)(k,j)    k ascending from 0, in ACC
 * (kgemm_nn.hpp:138-177, kgemm_nt.hpp:138-177, kgemm_tt.hpp:136-175), then
 *   beta == 1:  C += (T)(alpha*c)          atomic add (kgemm_nn.hpp:182-186)
 *   beta == 0:  C  = (T)(alpha*c)
 *   otherwise:  nn:    C = beta*C + (T)(alpha*c)  (alpha*c rounded to T first,
 *                                                  kgemm_nn.hpp:180, 191)
 *               nt/tt: C = beta*C + alpha*c       (kgemm_nt.hpp:192-194,
 *                                                  kgemm_tt.hpp:190-192)
 * with alpha and beta taken to ACC (kgemm_nn.hpp:32-33).
 * op = 0: nn, A mm x kk, B kk x nn;  op = 1: nt, B nn x kk;
 * op = 2: tt, A kk x mm, B nn x kk.  All column-major with per-entry leading
 * dimensions; C is mm x nn.
 */
#define DEFINE_KGEMM_ORACLE(T, ACC, SUF)                                       \
  void oracle_kgemm_batched_##SUF(int op, int mm, int nn, int kk, T alpha,     \
                                  const T *const *A, const int *ldA,           \
                                  const T *const *B, const int *ldB, T beta,   \
                                  T *const *C, const int *ldC, int batch)      \
  {                                                                            \
    const ACC al = alpha, be = beta;                                           \
    for (int b = 0; b < batch; b++)                                            \
    {                                                                          \
      const T *a = A[b], *bm = B[b];                                           \
      T *cm = C[b];                                                            \
      const size_t la = ldA[b], lb = ldB[b], lc = ldC[b];                      \
      for (int j = 0; j < nn; j++)                                             \
        for (int i = 0; i < mm; i++)                                           \
        {                                                                      \
          ACC c = 0;                                                           \
          for (int k = 0; k < kk; k++)                                         \
          {                                                                    \
            ACC x = (op == 2) ? a[k + i * la] : a[i + k * la];                 \
            ACC y = (op == 0) ? bm[k + j * lb] : bm[j + k * lb];               \
            c += x * y;                                                        \
          }                                                                    \
          T *cp = &cm[i + j * lc];                                             \
          if (be == 1)                                                         \
            *cp += (T)(al * c);                                                \
          else if (be == 0)                                                    \
            *cp = (T)(al * c);                                                 \
          else if (op == 0)                                                    \
          {                                                                    \
            const T acij = (T)(al * c);                                        \
            *cp = (T)(be * (ACC)(*cp) + (ACC)acij);                            \
          }                                                                    \
          else                                                                 \
            *cp = (T)(be * (ACC)(*cp) + al * c);                               \
        }                                                                      \
    }                                                                          \
  }

DEFINE_KGEMM_ORACLE(double, double, f64)
DEFINE_KGEMM_ORACLE(float, double, f32)

int oracle_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_num_threads(int t)
{
#ifdef _OPENMP
  if (t > 0)
    omp_set_num_threads(t);
#else
  (void)t;
#endif
}
