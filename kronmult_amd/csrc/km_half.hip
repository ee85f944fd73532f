// km_half.hip -- kron_half: the slab-streamed contraction split over two
// workgroups per entry by the rows of the last mode.
//
//   Y(i1, .., i_d) = sum_j1 A_1(i1, j1) [kron(A_2..A_d) X(j1, :)](i2, .., i_d)
//
// Workgroup `half` of a pair produces the outputs whose last index i_d lies in
// [half*N/2, (half+1)*N/2).  The backward order contracts mode d first, so
// splitting its rows splits every later stage too: the pair does exactly the
// work of one entry, with no stage repeated, and each workgroup's mode-1
// accumulator is half of one output (d6n6: 36 doubles per thread, against
// 96 in kron_slab, so 11 waves per CU instead of 8).  Per slab X(j1, :):
//   A  each thread loads its two mode-d fibers (N consecutive elements each)
//      straight into registers, one step ahead, and contracts them with its
//      half of A_d's rows into the half-slab in LDS;
//   B  modes d-1..2 are contracted in LDS, one fiber per thread per mode;
//   C  mode 1 is accumulated into registers: acc(i1, q) += A_1(i1, j1) Z(q).
// A step handles S slabs (1 or 2) side by side; the half-slabs are double
// buffered by step, so step u+1's stage A never waits for step u's stage C.  The two workgroups of a pair are 8 workgroups apart,
// which the dispatcher places on the same XCD: the second read of each slab
// is served by that XCD's L2.  The order of every dot product is the Alg. 993
// backward order (kronm_backward.hpp:92-139), mode 1 summed over j1
// ascending as in kron_slab.  FMA calls only (EXACT keeps kron_slab).
#include "km_internal.h"

namespace km
{

template<int D, int N>
struct HalfGeom
{
  static constexpr int H = N / 2;             // last-mode rows per workgroup
  static constexpr int FD = cpow(N, D - 2);   // mode-d fibers of one slab
  static constexpr int SL = FD * N;           // slab elements
  static constexpr int HS = FD * H;           // half-slab elements
  static constexpr int NT = HS / N;           // threads = fibers per LDS pass
  static_assert(N % 2 == 0, "the last mode splits in two equal halves");
  static_assert(FD == 2 * NT, "two mode-d fibers per thread");
};

// Modes D-2 .. 1 (0-based, i.e. d-1 .. 2) of the S half-slabs at hb,
// hb + HS, ..., whose layout is (i2, .., i_{d-1}, ii) with ii (the half's
// last-mode row) fastest: mode k has stride H * N^(D-2-k).  One fiber per
// thread and half-slab, S fibers per barrier.
template<typename T, int D, int N, int S, int K>
__device__ __forceinline__ void half_passes(T *hb, const EntryArgs<T> &a, int b,
                                            int tid)
{
  using C = typename Acc<T>::type;
  using G = HalfGeom<D, N>;
  constexpr int MODE = D - 2 - K;
  constexpr int STRIDE = G::H * cpow(N, D - 2 - MODE);
  C am[N][N];
  load_a<N, (N <= 6)>(am, a_of(a, b, MODE, D), a.lda);
  const int hi = tid / STRIDE;
  const int lo = tid - hi * STRIDE;
#pragma unroll
  for (int s = 0; s < S; s++)
    fiber_apply<N, false, T>(hb + s * G::HS + hi * STRIDE * N + lo, STRIDE, am);
  lds_barrier();
  if constexpr (MODE > 1)
    half_passes<T, D, N, S, K + 1>(hb, a, b, tid);
}

// One step handles S consecutive slabs j1 .. j1+S-1 of one entry (S divides
// N): their stages A and B run side by side, so each barrier separates S
// fibers of work per thread, and stage C adds them to the accumulator in j1
// order.
template<typename T, int D, int N, int S, bool PRELOAD>
__device__ __forceinline__ void half_body(const EntryArgs<T> &a, T *hbuf,
                                          int half, int p0, int p1)
{
  using C = typename Acc<T>::type;
  using G = HalfGeom<D, N>;
  constexpr int H = G::H, SL = G::SL, HS = G::HS, NT = G::NT;
  constexpr int RC = HS / NT; // accumulator positions per thread (= N)
  constexpr int SPE = N / S;  // steps per entry
  static_assert(N % S == 0, "steps cover whole entries");
  const int tid = threadIdx.x;

  C acc[N][RC];
  T *target = nullptr;
  // offset inside one output slab Y(i1, :) of accumulator position r
  auto yoff = [&](int r) {
    const int q = tid + NT * r;
    const int f = q / H;
    return f * N + half * H + (q - f * H);
  };
  auto flush = [&](T *dst) {
    if constexpr (PRELOAD)
    {
#pragma unroll
      for (int i1 = 0; i1 < N; i1++)
#pragma unroll
        for (int r = 0; r < RC; r++)
          gst(dst + i1 * SL + yoff(r), static_cast<T>(acc[i1][r]));
    }
    else
    {
#pragma unroll
      for (int i1 = 0; i1 < N; i1++)
#pragma unroll
        for (int r = 0; r < RC; r++)
          gatomic_add(dst + i1 * SL + yoff(r), static_cast<T>(acc[i1][r]));
    }
  };

  const int nstep = (p1 - p0) * SPE;
  // the thread's two mode-d fibers (N consecutive elements each) of every
  // slab of step u; pair loads when the vector allows them
  auto load = [&](T (&x)[S][2][N], int u) {
    const int p = p0 + u / SPE;
    const int j0 = (u - (u / SPE) * SPE) * S;
    const T *xe = x_of(a, entry_at(a, p));
    const bool al = reinterpret_cast<uintptr_t>(xe) % sizeof(pair_t<T>) == 0;
#pragma unroll
    for (int s = 0; s < S; s++)
#pragma unroll
      for (int f = 0; f < 2; f++)
      {
        const T *fp = xe + (long long)(j0 + s) * SL + (long long)(tid + f * NT) * N;
        if (al)
        {
#pragma unroll
          for (int j = 0; j < N; j += 2)
          {
            const pair_t<T> v = *(gptr<const pair_t<T>>)(fp + j);
            x[s][f][j] = v.x;
            x[s][f][j + 1] = v.y;
          }
        }
        else
        {
#pragma unroll
          for (int j = 0; j < N; j++)
            x[s][f][j] = gld(fp + j);
        }
      }
  };

  T x[S][2][N];
  load(x, 0);
  for (int u = 0; u < nstep; ++u)
  {
    const int p = p0 + u / SPE;
    const int j0 = (u - (u / SPE) * SPE) * S;
    const int b = entry_at(a, p);
    T *hb = hbuf + (u & 1) * S * HS;
    if (j0 == 0)
    {
      T *yb = y_of(a, b);
      if (yb != target)
      {
        if (target)
          flush(target);
#pragma unroll
        for (int i1 = 0; i1 < N; i1++)
#pragma unroll
          for (int r = 0; r < RC; r++)
            acc[i1][r] = PRELOAD ? static_cast<C>(gld(yb + i1 * SL + yoff(r))) : C(0);
        target = yb;
      }
    }
    // A: this half's rows of A_d on the step's fibers
    {
      const T *Ad = a_of(a, b, D - 1, D);
      C ad[H][N];
#pragma unroll
      for (int ii = 0; ii < H; ii++)
#pragma unroll
        for (int j = 0; j < N; j++)
          ad[ii][j] = static_cast<C>(cld(Ad + (half * H + ii) + j * a.lda));
#pragma unroll
      for (int s = 0; s < S; s++)
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
          for (int ii = 0; ii < H; ii++)
          {
            C v = ad[ii][0] * static_cast<C>(x[s][f][0]);
#pragma unroll
            for (int j = 1; j < N; j++)
              v = madd<false>(v, ad[ii][j], static_cast<C>(x[s][f][j]));
            hb[s * HS + (tid + f * NT) * H + ii] = static_cast<T>(v);
          }
    }
    // the registers are free: fetch the next step's fibers into them
    if (u + 1 < nstep)
      load(x, u + 1);
    lds_barrier();
    // B: modes d-1 .. 2
    half_passes<T, D, N, S, 0>(hb, a, b, tid);
    // C: mode 1 into the accumulator, slabs in j1 order
    const T *A1 = a_of(a, b, 0, D);
#pragma unroll
    for (int s = 0; s < S; s++)
    {
      C a1[N];
#pragma unroll
      for (int i1 = 0; i1 < N; i1++)
        a1[i1] = static_cast<C>(cld(A1 + (i1 + (j0 + s) * a.lda)));
#pragma unroll
      for (int r = 0; r < RC; r++)
      {
        const C z = static_cast<C>(hb[s * HS + tid + NT * r]);
#pragma unroll
        for (int i1 = 0; i1 < N; i1++)
          acc[i1][r] = madd<false>(acc[i1][r], a1[i1], z);
      }
    }
  }
  flush(target);
}

// Workgroup b serves pair ((b >> 4) << 3) | (b & 7), half (b >> 3) & 1: the
// two halves of a pair are 8 apart, on the same XCD.  The grid is a multiple
// of 16; workgroups of pairs >= npairs leave at once.
template<typename T, int D, int N, int S>
__global__ __launch_bounds__(cpow(N, D - 2) / 2) void kron_half(EntryArgs<T> a,
                                                               int npairs)
{
  plan_bind(a);
  using G = HalfGeom<D, N>;
  static_assert(G::NT == cpow(N, D - 2) / 2, "launch bounds");
  __shared__ T hbuf[2 * S * G::HS]; // two steps of S half-slabs
  const int b = blockIdx.x;
  const int half = (b >> 3) & 1;
  const int pair = ((b >> 4) << 3) | (b & 7);
  if (pair >= npairs)
    return;
  int p0, p1;
  if (a.ranges)
  {
    p0 = cld(a.ranges + pair);
    p1 = cld(a.ranges + pair + 1);
  }
  else
  {
    p0 = pair * a.chunk;
    p1 = min(a.batch, p0 + a.chunk);
  }
  if (p0 >= p1)
    return;
  // owned runs start from Y.  X keeps the default cache policy even when
  // streamed: each load instruction reads a third of every line it touches
  // (fibers of N elements per lane), and the rest of the line must come from
  // L1 for the next two instructions; non-temporal loads skip L1.
  if (a.owned)
    half_body<T, D, N, S, true>(a, hbuf, half, p0, p1);
  else
    half_body<T, D, N, S, false>(a, hbuf, half, p0, p1);
}

#define KM_HALF_CASE(T, D, N)                                                  \
  if (d == D && n == N)                                                        \
  {                                                                            \
    out->fn = (pf == 1) ? (const void *)&kron_half<T, D, N, 1>                 \
                        : (const void *)&kron_half<T, D, N, 2>;                \
    out->nt = HalfGeom<D, N>::NT;                                              \
    out->family = FAM_HALF;                                                    \
    return true;                                                               \
  }

template<>
bool half_kernel<double>(int d, int n, int pf, KernelInfo *out)
{
  KM_HALF_CASE(double, 6, 6)
  KM_HALF_CASE(double, 5, 8)
  return false;
}

template<>
bool half_kernel<float>(int d, int n, int pf, KernelInfo *out)
{
  KM_HALF_CASE(float, 6, 6)
  KM_HALF_CASE(float, 5, 8)
  return false;
}

} // namespace km
