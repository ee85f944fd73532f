// Batched slender GEMM: C_b = alpha*op(A_b)*op(B_b) + beta*C_b for every
// entry b, the contract of the reference's kgemm_{nn,nt,tt}_batched
// (kgemm_nn_batched.hpp:7-63 and the nt/tt copies; host wrappers
// kgemm_nn_batched.cpp:5-25).
//
// Every element is c = sum_k op(A)(i,k)*op(B)(k,j) over k ascending from
// zero in the accumulation type (double for float data), then
//   beta == 1: C += (T)(alpha*c) with a hardware atomic add (entries may
//              share C), kgemm_nn.hpp:182-186;
//   beta == 0: C  = (T)(alpha*c);
//   otherwise: C  = beta*C + alpha*c, where nn rounds alpha*c to T first
//              (kgemm_nn.hpp:180, 191) and nt/tt do not (kgemm_nt.hpp:192-194).
// EXACT multiplies and adds separately (the reference's CPU rounding); the
// default fuses them.
//
// Three kernels:
//   kgemm_rows: slender shapes (mm >= 64, kk <= 16).  A workgroup takes one
//     entry and a tile of 64, 128 or 256 rows; each thread holds
//     op(A)(i, :) in registers and walks the nn columns, reading op(B)(:, j)
//     through scalar loads (it is the same for the whole workgroup).  C is
//     written with consecutive rows on consecutive lanes.
//   kgemm_rows_small: mm < 64, kk <= 16.  The same per-row work over the
//     flattened (entry, row) space, so a workgroup holds many small entries.
//   kgemm_flat: kk > 16.  One thread per element of the flattened
//     (entry, j, i) space.
#include "km_internal.h"

namespace km
{
namespace
{

constexpr int kRowsNT = 256;
constexpr int kRowsKMax = 16;
constexpr int kFlatNT = 256;

template<int OP, typename T>
__device__ __forceinline__ const T *opa_ptr(const T *A, long long ld, int i, int k)
{
  // op(A)(i, k): A(i, k) for nn/nt, A(k, i) for tt
  return OP == 2 ? A + k + (long long)i * ld : A + i + (long long)k * ld;
}

template<int OP, typename T>
__device__ __forceinline__ const T *opb_ptr(const T *B, long long ld, int k, int j)
{
  // op(B)(k, j): B(k, j) for nn, B(j, k) for nt/tt
  return OP == 0 ? B + k + (long long)j * ld : B + j + (long long)k * ld;
}

// beta other than 0 and 1, with the old C value already loaded
template<int OP, typename T, typename C>
__device__ __forceinline__ T blend_c(T old, C c, C alpha, C beta)
{
  if (OP == 0)
  {
    const T acij = static_cast<T>(alpha * c);
    return static_cast<T>(beta * static_cast<C>(old) + static_cast<C>(acij));
  }
  return static_cast<T>(beta * static_cast<C>(old) + alpha * c);
}

// columns [j0, j0 + 8) of row i: for a general beta the eight old C values
// are loaded together before any is needed, so the row pays one memory
// latency per eight columns
template<int OP, bool EXACT, bool SCALAR_B, int KMAX, typename T, typename C>
__device__ __forceinline__ void row_columns(const C (&a)[KMAX], const T *B, long long lb,
                                            T *Cm, long long lc, int i, int j0,
                                            const KgemmArgs<T> &g)
{
  constexpr int JB = 8;
  const bool general = !(g.beta == C(1) || g.beta == C(0));
  T old[JB];
  if (general)
  {
#pragma unroll
    for (int u = 0; u < JB; u++)
      if (j0 + u < g.nn)
        old[u] = gld(Cm + i + (long long)(j0 + u) * lc);
  }
#pragma unroll
  for (int u = 0; u < JB; u++)
  {
    const int j = j0 + u;
    if (j >= g.nn)
      break;
    C c = 0;
#pragma unroll
    for (int k = 0; k < KMAX; k++)
      if (k < g.kk)
      {
        const T *bp = opb_ptr<OP>(B, lb, k, j);
        c = madd<EXACT>(c, a[k], static_cast<C>(SCALAR_B ? cld(bp) : gld(bp)));
      }
    T *cp = Cm + i + (long long)j * lc;
    if (general)
      gst(cp, blend_c<OP, T, C>(old[u], c, g.alpha, g.beta));
    else if (g.beta == C(1))
      gatomic_add(cp, static_cast<T>(g.alpha * c));
    else
      gst(cp, static_cast<T>(g.alpha * c));
  }
}

template<int OP, typename T, typename C>
__device__ __forceinline__ void store_c(T *cp, C c, C alpha, C beta)
{
  if (beta == C(1))
    gatomic_add(cp, static_cast<T>(alpha * c));
  else if (beta == C(0))
    gst(cp, static_cast<T>(alpha * c));
  else if (OP == 0)
  {
    const T acij = static_cast<T>(alpha * c);
    gst(cp, static_cast<T>(beta * static_cast<C>(gld(cp)) + static_cast<C>(acij)));
  }
  else
    gst(cp, static_cast<T>(beta * static_cast<C>(gld(cp)) + alpha * c));
}

template<typename T, int OP, bool EXACT>
__global__ __launch_bounds__(kRowsNT) void kgemm_rows(KgemmArgs<T> g)
{
  using C = typename Acc<T>::type;
  const long long total = (long long)g.batch * g.tiles;
  for (long long w = blockIdx.x; w < total; w += gridDim.x)
  {
    const int b = (int)(w / g.tiles);
    const int i = (int)(w - (long long)b * g.tiles) * blockDim.x + threadIdx.x;
    const T *A = cld(g.A + b);
    const T *B = cld(g.B + b);
    T *Cm = cld(g.C + b);
    const long long la = cld(g.lda + b), lb = cld(g.ldb + b), lc = cld(g.ldc + b);
    if (i >= g.mm)
      continue;
    C a[kRowsKMax];
#pragma unroll
    for (int k = 0; k < kRowsKMax; k++)
      if (k < g.kk)
        a[k] = static_cast<C>(gld(opa_ptr<OP>(A, la, i, k)));
    for (int j0 = 0; j0 < g.nn; j0 += 8)
      row_columns<OP, EXACT, true>(a, B, lb, Cm, lc, i, j0, g);
  }
}

// small matrices (mm < 64): one thread per row i of entry b over the
// flattened (b, i) space, so a workgroup holds many entries; op(A)(i, :) in
// registers, op(B) through vector loads (shared by the entry's threads in L1)
template<typename T, int OP, bool EXACT>
__global__ __launch_bounds__(kFlatNT) void kgemm_rows_small(KgemmArgs<T> g)
{
  using C = typename Acc<T>::type;
  const long long total = (long long)g.mm * g.batch;
  for (long long e = (long long)blockIdx.x * kFlatNT + threadIdx.x; e < total;
       e += (long long)gridDim.x * kFlatNT)
  {
    const int b = (int)(e / g.mm);
    const int i = (int)(e - (long long)b * g.mm);
    const T *A = gld(g.A + b);
    const T *B = gld(g.B + b);
    T *Cm = gld(g.C + b);
    const long long la = gld(g.lda + b), lb = gld(g.ldb + b), lc = gld(g.ldc + b);
    C a[kRowsKMax];
#pragma unroll
    for (int k = 0; k < kRowsKMax; k++)
      if (k < g.kk)
        a[k] = static_cast<C>(gld(opa_ptr<OP>(A, la, i, k)));
    for (int j0 = 0; j0 < g.nn; j0 += 8)
      row_columns<OP, EXACT, false>(a, B, lb, Cm, lc, i, j0, g);
  }
}

template<typename T, int OP, bool EXACT>
__global__ __launch_bounds__(kFlatNT) void kgemm_flat(KgemmArgs<T> g)
{
  using C = typename Acc<T>::type;
  const long long per = (long long)g.mm * g.nn;
  const long long total = per * g.batch;
  for (long long e = (long long)blockIdx.x * kFlatNT + threadIdx.x; e < total;
       e += (long long)gridDim.x * kFlatNT)
  {
    const int b = (int)(e / per);
    const long long r = e - (long long)b * per;
    const int j = (int)(r / g.mm);
    const int i = (int)(r - (long long)j * g.mm);
    const T *A = gld(g.A + b);
    const T *B = gld(g.B + b);
    T *Cm = gld(g.C + b);
    const long long la = gld(g.lda + b), lb = gld(g.ldb + b), lc = gld(g.ldc + b);
    C c = 0;
    for (int k = 0; k < g.kk; k++)
      c = madd<EXACT>(c, static_cast<C>(gld(opa_ptr<OP>(A, la, i, k))),
                      static_cast<C>(gld(opb_ptr<OP>(B, lb, k, j))));
    store_c<OP, T, C>(Cm + i + (long long)j * lc, c, g.alpha, g.beta);
  }
}

enum KgemmKind
{
  ROWS,
  ROWS_SMALL,
  FLAT
};

template<typename T, int OP>
const void *pick(KgemmKind k, bool exact)
{
  if (k == ROWS)
    return exact ? (const void *)&kgemm_rows<T, OP, true>
                 : (const void *)&kgemm_rows<T, OP, false>;
  if (k == ROWS_SMALL)
    return exact ? (const void *)&kgemm_rows_small<T, OP, true>
                 : (const void *)&kgemm_rows_small<T, OP, false>;
  return exact ? (const void *)&kgemm_flat<T, OP, true>
               : (const void *)&kgemm_flat<T, OP, false>;
}

} // namespace

template<typename T>
hipError_t kgemm_launch(int op, KgemmArgs<T> &g, bool exact, int cus,
                        hipStream_t stream)
{
  const KgemmKind kind = g.kk > kRowsKMax ? FLAT : (g.mm >= 64 ? ROWS : ROWS_SMALL);
  const void *fn = nullptr;
  switch (op)
  {
  case 0: fn = pick<T, 0>(kind, exact); break;
  case 1: fn = pick<T, 1>(kind, exact); break;
  default: fn = pick<T, 2>(kind, exact); break;
  }
  long long grid;
  int nt;
  if (kind == ROWS)
  {
    // row tiles of 64, 128 or 256 rows: no idle lanes for mm = 64 or 128
    nt = g.mm >= 256 ? kRowsNT : (g.mm >= 128 ? 128 : 64);
    g.tiles = (g.mm + nt - 1) / nt;
    grid = (long long)g.batch * g.tiles;
  }
  else if (kind == ROWS_SMALL)
  {
    g.tiles = 0;
    grid = ((long long)g.mm * g.batch + kFlatNT - 1) / kFlatNT;
    nt = kFlatNT;
  }
  else
  {
    g.tiles = 0;
    grid = ((long long)g.mm * g.nn * g.batch + kFlatNT - 1) / kFlatNT;
    nt = kFlatNT;
  }
  // grid-stride loops: cap the grid at many resident waves of workgroups
  const long long cap = (long long)(cus > 0 ? cus : 256) * 64;
  if (grid > cap)
    grid = cap;
  if (grid < 1)
    grid = 1;
  void *args[] = {&g};
  return hipLaunchKernel(fn, dim3((unsigned)grid), dim3(nt), args, 0, stream);
}

template hipError_t kgemm_launch<double>(int, KgemmArgs<double> &, bool, int,
                                         hipStream_t);
template hipError_t kgemm_launch<float>(int, KgemmArgs<float> &, bool, int,
                                        hipStream_t);

} // namespace km
