// km_xcd.hip -- kron_xcd: the two-pass contraction of kron_sub + kron_pair
// (d = 6, n = 7, 8: one n^(d-1) slab exceeds LDS) fused into one cooperative
// launch whose intermediate W stays in L2.
//
//   Z_b(j1, j2, :) = kron(A_3..A_d) X_b(j1, j2, :)              (pass 1)
//   Y(i1, i2, m)  += sum_j1 A_1(i1, j1) sum_j2 A_2(i2, j2) Z_b(j1, j2, m)
//                                                               (pass 2)
//
// Each XCD (8 on MI355X, 32 CUs and one 4 MB L2 each) takes a contiguous
// share of the order positions and runs its entries one after the other
// with all of its workgroups: pass 1 spreads the entry's n^2 sub-slabs over
// them and writes Z to an n^d buffer W private to the XCD (2 MiB at n = 8),
// an XCD-local barrier follows, then pass 2 spreads the n^(d-2) positions m
// over them.  W is written and read by workgroups of one XCD only, so it is
// served by that XCD's L2: HBM sees X, A and Y once.  The two-kernel form
// writes every W to HBM and reads it back (2.05x the compulsory bytes).
//
// Placement.  A workgroup reads the XCD it runs on from the XCC_ID hardware
// register and takes a rank among the workgroups of that XCD; the launch is
// cooperative (every workgroup resident at once), so the barriers cannot
// wait on a workgroup that never starts.  Every wait is bounded: a wait that
// expires (or an XCD with no workgroup) sets the launch's give-up word, every
// other wait then returns, every workgroup exits, and the sticky device
// counter read by km_cooperative_faults() records it.
//
// Memory.  Counters change only through device-scope atomics.  W is stored
// with ordinary stores (they write through L1 into the XCD's L2), each wave
// waits for its stores to complete before its workgroup arrives at the
// barrier, and pass 2 reads W with device-scope loads (`sc1`), which miss in
// L1: they see the current contents of the shared L2, never a line the L1
// kept from the previous entry.  Each workgroup stages its chunk of W in LDS
// with them, so every value crosses from L2 once (such loads reach 1.28 TB/s
// per XCD, against 2.17 for ordinary ones).  Workgroup-scope atomic loads
// (`sc0`) and ordinary loads after `buffer_inv sc0` may hit in L1 and gave
// stale values.  No device-scope fence is used: one would write W back to HBM
// and invalidate it in L2.
//
// Measured (d6n8 x1024 2:1): HBM fetch per call 2.18e9 B (X alone; the two
// kernels fetch 5.6e9), but 3.05 ms against 1.88 ms (atomic), and slower at
// every size tried, because the XCD's workgroups move through the phases in
// lockstep: per entry and XCD, pass 1 takes 8.1 us (X alone needs 3.4 at an
// eighth of the copy bandwidth), the barrier 2.2-3.2 us, pass 2 11 us (W
// staging 8 us of it).  So the variant table must select it (KM_KERNEL_XCD).
//
// Rounding: pass 1 and pass 2 are kron_sub's and kron_pair's FMA arithmetic
// (backward order, kronm_backward.hpp:92-139); pass 2 splits the mode-2 rows
// of a position over two threads, which changes no sum.  FMA calls only
// (KM_FLAG_EXACT keeps the two kernels).
#include "km_internal.h"

#include <atomic>
#include <map>
#include <mutex>

namespace km
{

namespace
{

constexpr int kXcdMax = 8;        // XCDs per GPU (MI355X)
// control block, in 32-bit words; counters 64 bytes apart
constexpr int XC_ARRIVED = 0;     // workgroups registered
constexpr int XC_GIVEUP = 16;     // nonzero: every wait returns false
constexpr int XC_RANK = 32;       // + 16*x: workgroups registered on XCD x
constexpr int XC_BAR = XC_RANK + 16 * 16; // + 16*x: barrier arrivals on XCD x
constexpr int XC_WORDS = XC_BAR + 16 * kXcdMax;
// ring of control blocks: consecutive launches use different blocks, so a
// launch's reset never races a still-running predecessor on another stream
// unless kXcdRing launches are in flight at once
constexpr int kXcdRing = 64;

__device__ unsigned km_xcd_ctl[kXcdRing * XC_WORDS];
__device__ unsigned long long km_xcd_faults;

__device__ __forceinline__ unsigned xc_add(unsigned *p, unsigned v)
{
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned xc_load(unsigned *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// one thread: wait until *p >= target; false once the launch gave up
__device__ bool xc_wait(unsigned *ctl, unsigned *p, unsigned target,
                        unsigned limit)
{
  for (unsigned it = 0;; ++it)
  {
    if (xc_load(p) >= target)
      return true;
    if (xc_load(ctl + XC_GIVEUP) != 0)
      return false;
    if (it >= limit)
    {
      if (xc_add(ctl + XC_GIVEUP, 1u) == 0)
        __hip_atomic_fetch_add(&km_xcd_faults, 1ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Whole workgroup: every wave's memory operations complete, then (thread 0)
// arrive at the barrier counter if `arrive`, and wait for `target` arrivals
// if target > 0.  Returns false once the launch gave up.
__device__ bool xc_sync(unsigned *ctl, unsigned *bar, bool arrive,
                        unsigned target, unsigned limit, int *s_ok)
{
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0)
  {
    if (arrive)
      xc_add(bar, 1u);
    *s_ok = target ? xc_wait(ctl, bar, target, limit) : 1;
  }
  __syncthreads();
  return *s_ok != 0;
}

__global__ void xcd_reset(unsigned *ctl)
{
  for (int i = threadIdx.x; i < XC_WORDS; i += blockDim.x)
    ctl[i] = 0;
}

} // namespace

// Q threads share one pass-2 position (each owns H of its mode-2 rows);
// the launch runs Q/2 workgroups per CU
template<typename T, int D, int N, int NT, int Q>
__global__ __launch_bounds__(NT, Q / 2) void kron_xcd(EntryArgs<T> a, unsigned *ctl,
                                               unsigned limit)
{
  plan_bind(a);
  using C = typename Acc<T>::type;
  constexpr int M = cpow(N, D - 2); // sub-slab elements = pass-2 positions
  constexpr int NN = N * N;         // sub-slabs per entry
  constexpr int R = (M + NT - 1) / NT;
  constexpr int MB = NT / Q;        // positions per pass-2 chunk
  constexpr int H = (N + Q - 1) / Q; // mode-2 rows per thread
  static_assert(MB % 64 == 0, "whole waves per row group");
  // pass 1's sub-slab, then pass 2's chunk of W
  constexpr int LB = M > NN * MB ? M : NN * MB;
  __shared__ __attribute__((aligned(16))) T buf[LB];
  T *const stage = buf;
  // A_1 and A_2 of the entry, row-major, for pass 2 (uniform LDS reads
  // instead of one scalar load per coefficient inside the j1 loop)
  __shared__ C sA[2 * N * N];
  __shared__ unsigned s_info[3];
  __shared__ int s_ok;
  const int tid = threadIdx.x;

  // ---- registration: XCD, rank on it, workgroups on it
  if (tid == 0)
  {
    const unsigned x = __builtin_amdgcn_s_getreg((31 << 11) | 20) & 0xFu;
    unsigned rank = 0;
    if (x < (unsigned)kXcdMax)
      rank = xc_add(ctl + XC_RANK + 16 * x, 1u);
    xc_add(ctl + XC_ARRIVED, 1u);
    bool ok = xc_wait(ctl, ctl + XC_ARRIVED, gridDim.x, limit);
    // every XCD must have a workgroup: its share of the entries is its own
    for (int y = 0; ok && y < kXcdMax; y++)
      ok = xc_load(ctl + XC_RANK + 16 * y) > 0;
    if (ok && x >= (unsigned)kXcdMax)
      ok = false;
    if (!ok && xc_add(ctl + XC_GIVEUP, 1u) == 0)
      __hip_atomic_fetch_add(&km_xcd_faults, 1ull, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    s_info[0] = x;
    s_info[1] = rank;
    s_info[2] = ok ? xc_load(ctl + XC_RANK + 16 * x) : 0;
  }
  __syncthreads();
  const int x = (int)s_info[0];
  const int rank = (int)s_info[1];
  const int P = (int)s_info[2];
  if (P == 0)
    return;
  unsigned *bar = ctl + XC_BAR + 16 * x;

  // ---- this XCD's order positions [lo, hi); owned runs never straddle two
  // XCDs (both ends move forward to the start of a run)
  const int batch = a.batch;
  int lo = (int)((long long)batch * x / kXcdMax);
  int hi = (int)((long long)batch * (x + 1) / kXcdMax);
  if (a.owned)
  {
    auto run_start = [&](int p) {
      while (p > 0 && p < batch && y_of(a, entry_at(a, p)) == y_of(a, entry_at(a, p - 1)))
        ++p;
      return p;
    };
    lo = run_start(lo);
    hi = run_start(hi);
  }
  if (lo >= hi)
    return;
  // the XCD's W: the workspace of its first entry (n^d elements)
  T *W = w_of(a, entry_at(a, lo), 0);

  // ---- pass-1 X prefetch, one sub-slab ahead, across entries
  T xr[R];
  auto load_x = [&](int p, int s) {
    const T *xp = x_of(a, entry_at(a, p)) + (long long)s * M;
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      const int q = tid + NT * r;
      if (M % NT == 0 || q < M)
        xr[r] = gld(xp + q);
    }
  };
  if (rank < NN)
    load_x(lo, rank);

  // ---- pass-2 geometry: chunks ch = rank, rank + P, .. of MB positions;
  // thread (mi, q) owns position ch*MB + mi and mode-2 rows [q*H, q*H + H)
  const int mi = tid % MB;
  // the row group is uniform across a wave: A_2's rows come through SGPRs
  const int q = __builtin_amdgcn_readfirstlane(tid / MB);
  const int i2lo = q * H;
  constexpr int NCH = (M + MB - 1) / MB;
  const int mych = rank < NCH ? (NCH - 1 - rank) / P + 1 : 0;
  // one chunk: runs of entries with one output sum in registers
  const bool single = mych == 1;
  C acc[N][H];
  T *target = nullptr;
  int tm = 0; // position of the accumulator (single chunk)

  auto acc_init = [&](T *yb, int m) {
    const bool live = m < M;
#pragma unroll
    for (int i1 = 0; i1 < N; i1++)
#pragma unroll
      for (int ii = 0; ii < H; ii++)
      {
        const int i2 = i2lo + ii;
        acc[i1][ii] = (a.owned && live && i2 < N)
                          ? static_cast<C>(gld(yb + (long long)(i1 * N + i2) * M + m))
                          : C(0);
      }
  };
  auto acc_flush = [&](T *yb, int m) {
    if (m >= M)
      return;
#pragma unroll
    for (int i1 = 0; i1 < N; i1++)
#pragma unroll
      for (int ii = 0; ii < H; ii++)
      {
        const int i2 = i2lo + ii;
        if (i2 < N)
        {
          T *yp = yb + (long long)(i1 * N + i2) * M + m;
          const T v = static_cast<T>(acc[i1][ii]);
          if (a.owned)
            gst(yp, v);
          else
            gatomic_add(yp, v);
        }
      }
  };

  unsigned k = 0; // barrier arrivals made by every workgroup of the XCD
  bool ok = true;
  for (int p = lo; p < hi && ok; ++p)
  {
    const int b = entry_at(a, p);
    // ---- pass 1: W may be overwritten once every workgroup has finished
    // the previous entry's pass 2 (its arrival k); the wait comes after the
    // first sub-slab's stages, so they overlap it
    bool freed = (p == lo);
    for (int s = rank; s < NN && ok; s += P)
    {
      lds_barrier();
#pragma unroll
      for (int r = 0; r < R; r++)
      {
        const int qq = tid + NT * r;
        if (M % NT == 0 || qq < M)
          buf[qq] = xr[r];
      }
      {
        int pn = p, sn = s + P;
        if (sn >= NN)
        {
          pn = p + 1;
          sn = rank;
        }
        if (pn < hi)
          load_x(pn, sn);
      }
      lds_barrier();
      trailing_stages<T, D, N, NT, false, 2, D - 1>(buf, a, b, tid);
      if (!freed)
      {
        ok = xc_sync(ctl, bar, false, k * P, limit, &s_ok);
        freed = true;
        if (!ok)
          break;
      }
      T *wb = W + (long long)s * M;
#pragma unroll
      for (int r = 0; r < R; r++)
      {
        const int qq = tid + NT * r;
        if (M % NT == 0 || qq < M)
          gst(wb + qq, buf[qq]);
      }
    }
    if (!ok)
      break;
    if (!freed)
    {
      ok = xc_sync(ctl, bar, false, k * P, limit, &s_ok);
      if (!ok)
        break;
    }
    // ---- W complete on this XCD
    ++k;
    ok = xc_sync(ctl, bar, true, k * P, limit, &s_ok);
    if (!ok)
      break;
    // ---- pass 2
    T *yb = y_of(a, b);
    const T *A1 = a_of(a, b, 0, D);
    const T *A2 = a_of(a, b, 1, D);
    for (int ch = rank; ch < NCH; ch += P)
    {
      const int m = ch * MB + mi;
      const bool live = m < M;
      if (single)
      {
        if (yb != target)
        {
          if (target)
            acc_flush(target, tm);
          acc_init(yb, m);
          target = yb;
          tm = m;
        }
      }
      else
        acc_init(yb, m);
      // the chunk's n^2 x MB values of W, each read from L2 once (device-
      // scope loads), into LDS; the Q threads of a position read them there
      __syncthreads(); // the previous chunk's (or pass 1's) readers are done
      for (int e = tid; e < 2 * NN; e += NT)
      {
        const int k = e / NN, r = e - k * NN; // k = 0: A_1, 1: A_2; r = i*N + j
        const int i = r / N, j = r - i * N;
        sA[e] = static_cast<C>(gld((k ? A2 : A1) + (i + j * a.lda)));
      }
#pragma unroll
      for (int e = tid; e < NN * MB; e += NT)
      {
        const int row = e / MB, em = ch * MB + (e - row * MB);
        if (em < M)
          stage[e] = __hip_atomic_load(W + (long long)row * M + em, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
      C w[N][N];
#pragma unroll
      for (int j1 = 0; j1 < N; j1++)
#pragma unroll
        for (int j2 = 0; j2 < N; j2++)
          w[j1][j2] = live ? static_cast<C>(stage[(j1 * N + j2) * MB + mi]) : C(0);
#pragma unroll
      for (int j1 = 0; j1 < N; j1++)
      {
        C u[H];
#pragma unroll
        for (int ii = 0; ii < H; ii++)
        {
          const int i2 = i2lo + ii;
          C t = 0;
          if (i2 < N)
          {
#pragma unroll
            for (int j2 = 0; j2 < N; j2++)
              t = madd<false>(t, sA[NN + i2 * N + j2], w[j1][j2]);
          }
          u[ii] = static_cast<C>(static_cast<T>(t));
        }
#pragma unroll
        for (int i1 = 0; i1 < N; i1++)
        {
          const C a1 = sA[i1 * N + j1];
#pragma unroll
          for (int ii = 0; ii < H; ii++)
            acc[i1][ii] = madd<false>(acc[i1][ii], a1, u[ii]);
        }
      }
      if (!single)
        acc_flush(yb, m);
    }
    // ---- arrival k+1: this workgroup has read W for the last time
    ++k;
    ok = xc_sync(ctl, bar, true, 0, limit, &s_ok);
  }
  if (ok && target)
    acc_flush(target, tm);
}

#ifndef KM_XCD_Q
#define KM_XCD_Q 4
#endif
#define KM_XCD_CASE(T, D, N, NT, Q)                                            \
  if (d == D && n == N)                                                        \
  {                                                                            \
    k->fn = (const void *)&kron_xcd<T, D, N, NT, Q>;                           \
    k->nt = NT;                                                                \
    k->family = FAM_XCD;                                                       \
    k->per_cu = Q / 2;                                                         \
    return true;                                                               \
  }

template<>
bool xcd_kernel<double>(int d, int n, KernelInfo *k)
{
  KM_XCD_CASE(double, 6, 7, 256, KM_XCD_Q)
  KM_XCD_CASE(double, 6, 8, 256, KM_XCD_Q)
  return false;
}

template<>
bool xcd_kernel<float>(int d, int n, KernelInfo *k)
{
  KM_XCD_CASE(float, 6, 7, 256, KM_XCD_Q)
  KM_XCD_CASE(float, 6, 8, 256, KM_XCD_Q)
  return false;
}

hipError_t xcd_prepare(hipStream_t st, unsigned **ctl)
{
  static std::mutex mu;
  static std::map<int, unsigned *> addr; // device -> its km_xcd_ctl
  static std::atomic<unsigned> next{0};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return e;
  unsigned *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it == addr.end())
    {
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(km_xcd_ctl));
      if (e != hipSuccess)
        return e;
      it = addr.emplace(dev, static_cast<unsigned *>(p)).first;
    }
    base = it->second;
  }
  unsigned *c = base + (size_t)(next.fetch_add(1) % kXcdRing) * XC_WORDS;
  // a fill kernel rather than a memset: capturable without memset nodes
  hipLaunchKernelGGL(xcd_reset, dim3(1), dim3(256), 0, st, c);
  e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  *ctl = c;
  return hipSuccess;
}

hipError_t xcd_faults(unsigned long long *count)
{
  return hipMemcpyFromSymbol(count, HIP_SYMBOL(km_xcd_faults), sizeof(*count), 0,
                             hipMemcpyDeviceToHost);
}

} // namespace km
