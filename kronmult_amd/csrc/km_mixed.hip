// km_mixed.hip -- matrix-core stages for the two-pass family at n = 8 (pass 1
// of kron_sub: modes d..3 of each (j1, j2) sub-slab, in LDS).
//
// A stage contracts mode k: out(i, f) = sum_j A_k(i, j) X(j, f) over the 512
// fibers f of a 4096-element sub-slab, K = n = 8.  Two arithmetics:
//
// KM_PRECISION_FP64_MFMA: v_mfma_f64_16x16x4f64 with M = the output row i,
//   N = 16 fibers, K = 4 (two instructions per K = 8).  Rows 8..15 of the tile
//   stay zero, so half of the issued multiply-adds are useful.  Lane l feeds
//   A(i = l%16, j = l/16 + 4h) and X(j = l/16 + 4h, fiber l%16), and receives
//   rows 4r + l/16 (r = 0..3) of fiber l%16: rows l/16 and l/16 + 4 are the
//   outputs (layout checked on the device).
//
// KM_PRECISION_F16X4 (mixed): v_mfma_f32_32x32x8f16 on integer slices, an
//   error-free split in the manner of the Ozaki scheme.  A fiber is scaled by
//   the power of two 2^-e_f that brings its largest element below 1, a row of
//   A_k by 2^-e_i; each scaled value v is cut into four 10-bit integer slices,
//   v = sum_s 2^(-10(s+1)) q_s + (rest below 2^-40), q_s in [-1023, 1023],
//   all exact in FP16.  One instruction takes the four slices of A_k (M = 4
//   slices x 8 rows = 32) against one slice of 32 fibers (N = 32, K = 8), so
//   four instructions give every slice pair.  Each product sum has at most
//   8 * 1023^2 < 2^23 in magnitude and is exact in FP32.  The 16 pair sums
//   are recombined in FP64 with their weights 2^(-10(s+t+2)) and the scales:
//   the stage result carries the 40-bit slicing error of X and of A_k
//   (relative to the fiber's and the row's largest element), nothing else.
//   Pass 2 (modes 2 and 1, the accumulation into Y) stays FP64.
// No counterpart in the reference (its stages are kgemm_tt, FP64 scalar
// loops, kgemm_tt.hpp:118-194).
#include "km_internal.h"

namespace km
{

typedef double v4d __attribute__((ext_vector_type(4)));
typedef _Float16 v4h __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

// FP64 MFMA stage of one mode: STRIDE is the mode's stride in the sub-slab,
// FPW fibers per wavefront
template<int N, int FPW, int STRIDE>
__device__ __forceinline__ void mx_stage_f64(double *buf, const double *Ak, int lda,
                                             int tid)
{
  static_assert(N == 8 && FPW % 16 == 0, "K = 8 in two 16x16x4 steps");
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, c = lane & 15;
  double av[2];
#pragma unroll
  for (int h = 0; h < 2; h++)
    av[h] = (c < N) ? gld(Ak + c + (g + 4 * h) * lda) : 0.0;
#pragma unroll
  for (int grp = 0; grp < FPW / 16; grp++)
  {
    const int f = wave * FPW + grp * 16 + c;
    const int hi = f / STRIDE, lo = f - hi * STRIDE;
    double *fp = buf + hi * STRIDE * N + lo;
    const double b0 = fp[g * STRIDE];
    const double b1 = fp[(g + 4) * STRIDE];
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[0], b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[1], b1, acc, 0, 0, 0);
    fp[g * STRIDE] = acc[0];
    fp[(g + 4) * STRIDE] = acc[1];
  }
}

// exponent e with |v| < 2^e for the largest |v| (0 for zero)
__device__ __forceinline__ int mx_exp(double m)
{
  int e = 0;
  (void)__builtin_frexp(m, &e);
  return m == 0.0 ? 0 : e;
}

// four 10-bit integer slices of v (|v| < 1), as FP16
__device__ __forceinline__ void mx_slices(double v, _Float16 (&q)[4])
{
#pragma unroll
  for (int s = 0; s < 4; s++)
  {
    v *= 1024.0;
    const double t = __builtin_trunc(v);
    v -= t;
    q[s] = static_cast<_Float16>(static_cast<float>(t));
  }
}

// mixed stage of one mode (see the file comment)
template<int N, int FPW, int STRIDE>
__device__ __forceinline__ void mx_stage_f16x4(double *buf, const double *Ak, int lda,
                                               int tid)
{
  static_assert(N == 8 && FPW % 32 == 0, "K = 8, 32 fibers per instruction");
  const int lane = tid & 63, wave = tid >> 6;
  const int kh = lane >> 5, c = lane & 31;
  // A operand: row m = c of the 32 x 8 tile is slice t = c/8 of row i = c%8
  const int ta = c >> 3, ia = c & 7;
  double arow[N];
  double amax = 0.0;
#pragma unroll
  for (int j = 0; j < N; j++)
  {
    arow[j] = gld(Ak + ia + j * lda);
    amax = fmax(amax, fabs(arow[j]));
  }
  const int ea = mx_exp(amax);
  v4h aop;
#pragma unroll
  for (int q = 0; q < 4; q++)
  {
    _Float16 sl[4];
    mx_slices(__builtin_ldexp(arow[4 * kh + q], -ea), sl);
    aop[q] = ta == 0 ? sl[0] : ta == 1 ? sl[1] : ta == 2 ? sl[2] : sl[3];
  }
  // exponents of the output rows 4*kh + rr (lane i holds row i's)
  int erow[4];
#pragma unroll
  for (int rr = 0; rr < 4; rr++)
    erow[rr] = __shfl(ea, 4 * kh + rr, 64);
  const v16f zero = {};
#pragma unroll 1
  for (int grp = 0; grp < FPW / 32; grp++)
  {
    const int f = wave * FPW + grp * 32 + c;
    const int hi = f / STRIDE, lo = f - hi * STRIDE;
    double *fp = buf + hi * STRIDE * N + lo;
    double x[4];
    double xm = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++)
    {
      x[q] = fp[(4 * kh + q) * STRIDE];
      xm = fmax(xm, fabs(x[q]));
    }
    int ef = mx_exp(xm);
    ef = max(ef, __shfl_xor(ef, 32, 64));
    v4h bop[4];
#pragma unroll
    for (int q = 0; q < 4; q++)
    {
      _Float16 sl[4];
      mx_slices(__builtin_ldexp(x[q], -ef), sl);
#pragma unroll
      for (int s = 0; s < 4; s++)
        bop[s][q] = sl[s];
    }
    v16f d[4];
#pragma unroll
    for (int s = 0; s < 4; s++)
      d[s] = __builtin_amdgcn_mfma_f32_32x32x8f16(aop, bop[s], zero, 0, 0, 0);
    // lane holds rows m = 8*(r/4) + 4*kh + r%4: slice t = r/4 of row 4*kh + r%4
#pragma unroll
    for (int rr = 0; rr < 4; rr++)
    {
      double sc[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int t = 0; t < 4; t++)
          sc[s + t] += static_cast<double>(d[s][4 * t + rr]);
      double acc = sc[6];
#pragma unroll
      for (int k = 5; k >= 0; k--)
        acc = acc * (1.0 / 1024.0) + sc[k];
      fp[(4 * kh + rr) * STRIDE] = __builtin_ldexp(acc, ef + erow[rr] - 20);
    }
  }
}

// modes MODE, MODE-1, ..., 2 (0-based) of the sub-slab, one barrier each
template<int D, int N, int NT, int PREC, int MODE>
__device__ __forceinline__ void mx_stages(double *buf, const EntryArgs<double> &a,
                                          int b, int tid)
{
  constexpr int M = cpow(N, D - 2);
  constexpr int FPW = M / N / (NT / 64);
  constexpr int STRIDE = cpow(N, D - 1 - MODE);
  const double *Ak = a_of(a, b, MODE, D);
  if constexpr (PREC == KM_PRECISION_FP64_MFMA)
    mx_stage_f64<N, FPW, STRIDE>(buf, Ak, a.lda, tid);
  else
    mx_stage_f16x4<N, FPW, STRIDE>(buf, Ak, a.lda, tid);
  lds_barrier();
  if constexpr (MODE > 2)
    mx_stages<D, N, NT, PREC, MODE - 1>(buf, a, b, tid);
}

// pass 1 with matrix-core stages: as kron_sub (sub-entries s in [0, nsub)
// of the chunk starting at order position e0)
template<int D, int N, int NT, int PREC>
__global__ __launch_bounds__(NT) void kron_sub_mx(EntryArgs<double> a, int e0,
                                                  int nsub)
{
  constexpr int M = cpow(N, D - 2);
  constexpr int R = M / NT;
  constexpr int NN = N * N;
  static_assert(M % NT == 0, "whole sub-slab per thread count");
  __shared__ double buf[M];
  const int tid = threadIdx.x;
  const int s0 = blockIdx.x * a.chunk;
  const int s1 = min(nsub, s0 + a.chunk);
  if (s0 >= s1)
    return;
  double xr[R];
  {
    const int b = entry_at(a, e0 + s0 / NN);
    const double *xp = x_of(a, b) + (long long)(s0 % NN) * M;
#pragma unroll
    for (int r = 0; r < R; r++)
      xr[r] = gld(xp + tid + NT * r);
  }
  for (int s = s0; s < s1; ++s)
  {
    const int p = e0 + s / NN;
    const int js = s % NN;
    const int b = entry_at(a, p);
    lds_barrier();
#pragma unroll
    for (int r = 0; r < R; r++)
      buf[tid + NT * r] = xr[r];
    if (s + 1 < s1)
    {
      const int bn = entry_at(a, e0 + (s + 1) / NN);
      const double *xp = x_of(a, bn) + (long long)((s + 1) % NN) * M;
#pragma unroll
      for (int r = 0; r < R; r++)
        xr[r] = gld(xp + tid + NT * r);
    }
    lds_barrier();
    mx_stages<D, N, NT, PREC, D - 1>(buf, a, b, tid);
    double *wb = w_of(a, b, p - e0) + (long long)js * M;
#pragma unroll
    for (int r = 0; r < R; r++)
      gst(wb + tid + NT * r, buf[tid + NT * r]);
  }
}

template<>
bool twopass_mx_kernel<double>(int d, int n, int precision, KernelInfo *sub)
{
  if (d == 6 && n == 8)
  {
    if (precision == KM_PRECISION_FP64_MFMA)
      sub->fn = (const void *)&kron_sub_mx<6, 8, 256, KM_PRECISION_FP64_MFMA>;
    else if (precision == KM_PRECISION_F16X4)
      sub->fn = (const void *)&kron_sub_mx<6, 8, 256, KM_PRECISION_F16X4>;
    else
      return false;
    sub->nt = 256;
    sub->family = FAM_TWOPASS;
    return true;
  }
  return false;
}

template<>
bool twopass_mx_kernel<float>(int, int, int, KernelInfo *)
{
  return false;
}

} // namespace km
