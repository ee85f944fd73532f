// km_device.h -- gfx950 device code for the batched Kronecker mat-vec
//     Y(:,b) += kron(A_1(b), ..., A_d(b)) * X(:,b),   b = 0..batch-1.
//
// Index convention (the reference's, test_kronmult6_batched.cpp:352-385):
// X and Y are column-major tensors with j_d fastest and j_1 slowest; A_k is
// column-major, A_k(i,j) = Ak[i + j*lda] (kgemm_nt.hpp:81).  Mode k (0-based)
// of a tensor of n^d elements therefore has stride n^(d-1-k).
//
// Kernel families:
//   kron_lds   whole n^d tensor of the entry in LDS; every mode contraction is
//              a pass of per-thread fibers (n loads, n*n FMAs, n stores);
//   kron_reg   n^(d-RM) threads of n^RM values each: groups of modes
//              contracted in registers, one LDS transpose between groups;
//   kron_slab  n^d too large for LDS but n^(d-1) is not: the input is streamed
//              in slabs X(j1,:) (direct-to-LDS loads, double buffered), modes
//              2..d are contracted inside the slab and mode 1 is accumulated
//              straight into per-thread registers;
//   kron_sub + kron_pair  even one slab too large (d = 6, n = 7, 8): two
//              passes through the caller's workspace;
//   kron_runs  entries of different n in one launch, whole output runs per
//              workgroup (km_runs.hip);
//   kron_global  any shape; stages through the caller's workspace in global
//              memory (the reference's own storage scheme) -- the fallback.
// The first three need no global workspace.
//
// Output handling shared by all families: a workgroup walks a contiguous range
// of entries (in plan order when a plan is given).  Consecutive entries that
// target the same Y are summed on chip; a run is flushed with one 64-bit
// atomic add per element (or a plain read-modify-write when the plan
// guarantees that the workgroup owns every entry of that output).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace km
{

// largest mode extent n_k whose fiber the workspace-staged kernel holds in
// registers; wider fibers go to dynamic LDS (KM_GLOBAL_LDS_BYTES per
// workgroup, which bounds n_k * workgroup size)
#define KM_GLOBAL_MAX_N 32
#define KM_GLOBAL_LDS_BYTES (150 * 1024)

// fiber-loop unroll factors of the LDS stage passes (tuning knobs)
#ifndef KM_LDS_UNROLL
#define KM_LDS_UNROLL 4
#endif
#ifndef KM_SLAB_UNROLL
#define KM_SLAB_UNROLL 1
#endif
constexpr int kLdsUnroll = KM_LDS_UNROLL;
constexpr int kSlabUnroll = KM_SLAB_UNROLL;


// ---------------------------------------------------------------------------
// arguments
// ---------------------------------------------------------------------------
template<typename T>
struct EntryArgs
{
  // A_k of entry b: Aptr ? Aptr[k + d*b] : A + b*a_stride + k*a_kstride
  const T *A;
  const T *const *Aptr;
  long long a_stride;
  long long a_kstride;
  int lda;
  // X of entry b: pX ? pX[b] : X + b*x_stride
  const T *X;
  const T *const *pX;
  long long x_stride;
  // Y of entry b: pY ? pY[b] : Y + b*y_stride
  T *Y;
  T *const *pY;
  long long y_stride;
  // workspace (kron_global only): slot s buffers are W0(s), W1(s)
  T *const *pW;       // pW[ws_mul*s + 0], pW[ws_mul*s + 1]
  T *W;               // W + s*w_stride, W + s*w_stride + w_half
  long long w_stride;
  long long w_half;
  int ws_mul;         // 1: pW[s] (one buffer per slot), 2: pW pairs
  int batch;
  int chunk;          // entries per workgroup when ranges == nullptr
  const int *ranges;  // workgroup w walks order positions [ranges[w], ranges[w+1])
  const int *order;   // order position -> entry
  int owned;          // 1: workgroup owns its outputs (no atomics)
  int x_stream;       // 1: every X is read once per call (non-temporal loads)
  int slot_by_entry;  // two-pass family: workspace slot = entry index (1) or
                      // position inside the launched chunk (0)
  // plan binding: when *bind_flag == bind_gen the caller's pY no longer
  // matches the pointers the plan was built from, and the call accumulates
  // with atomics instead of the plan's owned read-modify-writes
  const int *bind_flag;
  unsigned bind_gen;
};

struct Shape
{
  int d;
  int m[6];
  int n[6];
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
constexpr int cpow(int b, int e) { return e == 0 ? 1 : b * cpow(b, e - 1); }

// Address-space views.  Pointer tables, plan arrays and A are read-only for
// the whole launch, so they are read through the constant address space
// (scalar loads into SGPRs when the address is wave-uniform); X, Y and the
// workspace are read and written through the global address space.
template<typename T>
using cptr = const __attribute__((address_space(4))) T *;
template<typename T>
using gptr = __attribute__((address_space(1))) T *;

template<typename T>
__device__ __forceinline__ T cld(const T *p)
{
  return *(cptr<T>)(p);
}

template<typename T>
__device__ __forceinline__ T gld(const T *p)
{
  return *(gptr<const T>)(p);
}

template<typename T>
__device__ __forceinline__ void gst(T *p, T v)
{
  *(gptr<T>)(p) = v;
}

template<typename T>
__device__ __forceinline__ void gatomic_add(T *p, T v)
{
  __hip_atomic_fetch_add((gptr<T>)(p), v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

template<typename T>
__device__ __forceinline__ void wg_range(const EntryArgs<T> &a, int &p0, int &p1)
{
  if (a.ranges)
  {
    p0 = cld(a.ranges + blockIdx.x);
    p1 = cld(a.ranges + blockIdx.x + 1);
  }
  else
  {
    p0 = blockIdx.x * a.chunk;
    p1 = min(a.batch, p0 + a.chunk);
  }
}

// First statement of every kernel that flushes outputs: a plan whose recorded
// outputs differ from the caller's pY (found by the check kernel launched
// just before) gives up ownership, so every flush becomes an atomic add into
// the caller's outputs -- correct for any pY, never a race.
template<typename T>
__device__ __forceinline__ void plan_bind(EntryArgs<T> &a)
{
  if (a.bind_flag && (unsigned)cld(a.bind_flag) == a.bind_gen)
    a.owned = 0;
}

template<typename T>
__device__ __forceinline__ int entry_at(const EntryArgs<T> &a, int p)
{
  return a.order ? cld(a.order + p) : p;
}

// X element load.  When every X vector is read once per call (batched,
// pbatched, vbatched, grouped: EntryArgs::x_stream) the load carries the
// non-temporal hint, so the streamed vectors do not push reused lines (A, the
// next entries' prefetched A, Y) out of L2; pooled X (xbatched) keeps the
// default policy, since entries share vectors there.  A compile-time choice:
// the compiler merges a runtime select of the two loads into one plain load.
template<bool XNT, typename T>
__device__ __forceinline__ T gld_x(const T *p)
{
  if constexpr (XNT)
    return __builtin_nontemporal_load((gptr<const T>)(p));
  else
    return *(gptr<const T>)(p);
}

template<typename T>
__device__ __forceinline__ const T *x_of(const EntryArgs<T> &a, int b)
{
  return a.pX ? cld(a.pX + b) : a.X + (long long)b * a.x_stride;
}

template<typename T>
__device__ __forceinline__ T *y_of(const EntryArgs<T> &a, int b)
{
  return a.pY ? cld(a.pY + b) : a.Y + (long long)b * a.y_stride;
}

template<typename T>
__device__ __forceinline__ const T *a_of(const EntryArgs<T> &a, int b, int k,
                                         int d)
{
  return a.Aptr ? cld(a.Aptr + (k + d * b))
                : a.A + (long long)b * a.a_stride + k * a.a_kstride;
}

// acc += a*x.  EXACT: one rounded multiply then one rounded add, the
// reference's CPU rounding (kgemm_nt.hpp:125-134; this library is compiled
// with -ffp-contract=off so nothing fuses them).  Otherwise a fused FMA.
template<bool EXACT, typename T>
__device__ __forceinline__ T madd(T acc, T a, T x)
{
  if constexpr (EXACT)
  {
    T p = a * x;
    return acc + p;
  }
  else
  {
    return __builtin_fma(a, x, acc);
  }
}

// Float entry points accumulate each dot product in double and round the
// stored result to float, as the reference's float templates do
// (kgemm_nt.hpp:213-221).
template<typename T>
struct Acc
{
  using type = T;
};
template<>
struct Acc<float>
{
  using type = double;
};

// A wave-uniform double held in VGPRs (a float widened by the vector unit)
// moved to an SGPR pair with the readfirstlane builtin.  The compiler folds
// a readfirstlane of a value it knows to be uniform (and then keeps the
// double in two VGPRs per lane), so the bits are first OR-ed with mbcnt(0),
// which is 0 in every lane but not known to be uniform.  The builtin, unlike
// inline asm, gets the wait states the result of the widening needs.
__device__ __forceinline__ double to_sgpr(double v)
{
  const unsigned z = __builtin_amdgcn_mbcnt_lo(0u, 0u);
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u | z);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32) | z);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Load A_k (n x n, leading dimension lda) into registers.  The pointer is
// wave-uniform, so these become scalar loads and the values live in SGPRs.
// Float factors are widened by the vector unit and stay in VGPRs unless
// SGPR_WIDE moves them back to SGPRs: kron_slab at n <= 6 needs that (its
// double accumulator leaves no room for 2*n*n more VGPRs; float d6n6 1.12 ->
// 0.86 ms); the other kernels, and n = 8, measured slower with it.
template<int N, bool SGPR_WIDE = false, typename T>
__device__ __forceinline__ void load_a(typename Acc<T>::type (&am)[N][N],
                                       const T *Ak, int lda)
{
  using C = typename Acc<T>::type;
#pragma unroll
  for (int j = 0; j < N; j++)
#pragma unroll
    for (int i = 0; i < N; i++)
    {
      const C v = static_cast<C>(cld(Ak + (i + j * lda)));
      if constexpr (SGPR_WIDE && sizeof(C) != sizeof(T))
        am[i][j] = to_sgpr(v);
      else
        am[i][j] = v;
    }
}

// out_i = sum_j am[i][j] * x_j over j ascending (started from zero as in the
// reference), written back over the fiber's own n slots.
template<int N, bool EXACT, typename T>
__device__ __forceinline__ void fiber_apply(T *p, int s,
                                            const typename Acc<T>::type (&am)[N][N])
{
  using C = typename Acc<T>::type;
  C x[N];
#pragma unroll
  for (int j = 0; j < N; j++)
    x[j] = static_cast<C>(p[j * s]);
#pragma unroll
  for (int i = 0; i < N; i++)
  {
    C acc = 0;
    if constexpr (EXACT)
    {
#pragma unroll
      for (int j = 0; j < N; j++)
        acc = madd<true>(acc, am[i][j], x[j]);
    }
    else
    {
      acc = am[i][0] * x[0];
#pragma unroll
      for (int j = 1; j < N; j++)
        acc = madd<false>(acc, am[i][j], x[j]);
    }
    p[i * s] = static_cast<T>(acc);
  }
}

template<typename T>
__device__ __forceinline__ void flush_add(T *dst, T v, int owned)
{
  if (owned)
    gst(dst, gld(dst) + v);
  else
    gatomic_add(dst, v);
}

// dst[tid + NT*r] += acc[r] for r < R (positions below S).  Owned outputs are
// read-modified-written with all loads of a chunk issued before its stores,
// so the memory latency is paid once per chunk rather than once per element;
// shared outputs use one hardware FP64 atomic add per element.
template<int R, int NT, int S, int CHUNK, typename T, typename C>
__device__ __forceinline__ void flush_run(T *dst, const C (&acc)[R], int tid,
                                          int owned)
{
  if (owned)
  {
#pragma unroll
    for (int r0 = 0; r0 < R; r0 += CHUNK)
    {
      T yv[CHUNK];
#pragma unroll
      for (int c = 0; c < CHUNK; c++)
      {
        const int r = r0 + c, q = tid + NT * r;
        if (r < R && (S % NT == 0 || q < S))
          yv[c] = gld(dst + q);
      }
#pragma unroll
      for (int c = 0; c < CHUNK; c++)
      {
        const int r = r0 + c, q = tid + NT * r;
        if (r < R && (S % NT == 0 || q < S))
          gst(dst + q, yv[c] + static_cast<T>(acc[r]));
      }
    }
  }
  else
  {
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      const int q = tid + NT * r;
      if (S % NT == 0 || q < S)
        gatomic_add(dst + q, static_cast<T>(acc[r]));
    }
  }
}

// Output runs (consecutive entries with the same output, summed on chip).  A
// run whose output the workgroup owns (a plan) starts from Y itself when the
// contraction uses FMAs: the loads are issued at the start of the run and
// are not waited for until the first update, after the entry's mode passes,
// and the end of the run is a plain store.  EXACT keeps the reference's
// rounding Y + (sum), so it starts from zero and reads Y back at the end;
// shared outputs use atomics.
template<int R, int NT, int S, typename T, typename C>
__device__ __forceinline__ void run_init(C (&acc)[R], const T *src, int tid,
                                         bool preload)
{
#pragma unroll
  for (int r = 0; r < R; r++)
  {
    const int q = tid + NT * r;
    acc[r] = (preload && (S % NT == 0 || q < S)) ? static_cast<C>(gld(src + q)) : C(0);
  }
}

template<int R, int NT, int S, int CHUNK, typename T, typename C>
__device__ __forceinline__ void run_flush(T *dst, const C (&acc)[R], int tid,
                                          int owned, bool preloaded)
{
  if (preloaded)
  {
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      const int q = tid + NT * r;
      if (S % NT == 0 || q < S)
        gst(dst + q, static_cast<T>(acc[r]));
    }
  }
  else
  {
    flush_run<R, NT, S, CHUNK>(dst, acc, tid, owned);
  }
}

// Pull the next entry's A matrices into L2 with one vector load per 128-byte
// line; the value is consumed (keep_alive) only after the current entry, so
// the later scalar loads of that entry hit L2 instead of HBM.
template<int D, int N, typename T>
__device__ __forceinline__ T prefetch_a(const EntryArgs<T> &a, int b, int tid)
{
  constexpr int PER_LINE = 128 / (int)sizeof(T);
  T v = 0;
  const int span = (N - 1) * (a.lda ? a.lda : N) + N; // elements of one matrix
  const int lines = (span + PER_LINE - 1) / PER_LINE;
#pragma unroll
  for (int k = 0; k < D; k++)
  {
    const T *Ak = a_of(a, b, k, D);
    if (tid < lines)
      v += gld(Ak + min(tid * PER_LINE, span - 1));
  }
  return v;
}

// Whether kron_lds / kron_reg pull the next entry's A into L2 during the
// current entry.  Measured: it pays for 7x7 and 8x8 factors (d2n8 1.43 ->
// 1.10 ms, d3n8 0.53 -> 0.49 ms, d5n7 0.223 -> 0.207 ms) and costs for n <= 6,
// whose few lines are cheaper to load on demand (d6n4 0.64 -> 0.57 ms,
// d3n4 0.83 -> 0.73 ms, d6n3 0.35 -> 0.31 ms without it).
template<int N>
constexpr bool kPrefetchA = N >= 7;

template<typename T>
__device__ __forceinline__ void keep_alive(T v)
{
  asm volatile("" ::"v"(v));
}

// ---------------------------------------------------------------------------
// kron_lds: whole tensor in LDS, one workgroup per entry at a time.
//   FWD = false: modes contracted d..1 (Alg. 993 backward order,
//                kronm_backward.hpp:92-139)
//   FWD = true : modes contracted 1..d (the recursive / forward order,
//                kronmultv_recur.hpp:90-145, kronm_forward.hpp:110-164)
// ---------------------------------------------------------------------------
// Barrier between mode passes: the whole workgroup, or (WAVE) only the
// wavefront that owns the buffer.  LDS traffic of one wavefront is performed
// in order; the workgroup-scope fence waits for it and keeps the compiler
// from moving LDS accesses across the barrier.
// s_waitcnt immediate for gfx9: lgkmcnt(0), vmcnt and expcnt left at their
// maximum (not waited on)
constexpr int kWaitLgkm0 = 0xC07F;

// Workgroup barrier for data exchanged through LDS only.  It waits for this
// wave's LDS (and scalar) accesses, then for every wave, but not for vector
// memory: global loads in flight (the next entry's X prefetched into
// registers, the slab DMA into the other LDS buffer) keep running across it.
// __syncthreads() would wait for them (on gfx9 vmcnt counts loads as well as
// stores).  The empty asm statements stop the compiler from moving memory
// accesses across the barrier.  This is no fence for global memory; the
// kernels that use it never pass data between threads through global memory.
__device__ __forceinline__ void lds_barrier()
{
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

template<bool WAVE>
__device__ __forceinline__ void pass_sync()
{
  if constexpr (WAVE)
  {
    // one wavefront: its LDS accesses complete in order once lgkmcnt is 0
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
  }
  else
  {
    lds_barrier();
  }
}

template<typename T, int D, int N, int NT, bool EXACT, bool FWD, int K,
         bool WAVE = false>
__device__ __forceinline__ void lds_stage(T *buf, const EntryArgs<T> &a, int b,
                                          int tid)
{
  using C = typename Acc<T>::type;
  constexpr int S = cpow(N, D);
  constexpr int F = S / N;
  constexpr int MODE = FWD ? K : D - 1 - K;
  constexpr int STRIDE = cpow(N, D - 1 - MODE);
  C am[N][N];
  // lda 0: each A_k is stored densely (leading dimension n), as in the
  // grouped entry point where n differs from entry to entry
  load_a<N>(am, a_of(a, b, MODE, D), a.lda ? a.lda : N);
#pragma unroll kLdsUnroll
  for (int f = tid; f < F; f += NT)
  {
    int hi = f / STRIDE;
    int lo = f - hi * STRIDE;
    fiber_apply<N, EXACT, T>(buf + hi * STRIDE * N + lo, STRIDE, am);
  }
  pass_sync<WAVE>();
  if constexpr (K + 1 < D)
    lds_stage<T, D, N, NT, EXACT, FWD, K + 1, WAVE>(buf, a, b, tid);
}

template<typename T, int D, int N, int NT, bool EXACT, bool FWD, bool PRELOAD,
         bool XNT>
__device__ __forceinline__ void lds_body(const EntryArgs<T> &a, T *buf, int p0,
                                         int p1)
{
  constexpr int S = cpow(N, D);
  constexpr int R = (S + NT - 1) / NT;
  const int tid = threadIdx.x;

  T acc[R];
  T xr[R];
  T *target = nullptr;
  constexpr bool preload = PRELOAD;

  {
    const T *xp = x_of(a, entry_at(a, p0));
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      int q = tid + NT * r;
      if (S % NT == 0 || q < S)
        xr[r] = gld_x<XNT>(xp + q);
    }
  }

  for (int p = p0; p < p1; ++p)
  {
    const int b = entry_at(a, p);
    T *yb = y_of(a, b);
    lds_barrier(); // buf free: previous entry fully consumed
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      int q = tid + NT * r;
      if (S % NT == 0 || q < S)
        buf[q] = xr[r];
    }
    if (p + 1 < p1)
    {
      const T *xp = x_of(a, entry_at(a, p + 1));
#pragma unroll
      for (int r = 0; r < R; r++)
      {
        int q = tid + NT * r;
        if (S % NT == 0 || q < S)
          xr[r] = gld_x<XNT>(xp + q);
      }
    }
    T pf = 0;
    if (kPrefetchA<N> && p + 1 < p1)
      pf = prefetch_a<D, N>(a, entry_at(a, p + 1), tid);
    if (yb != target)
    {
      if (target)
        run_flush<R, NT, S, 8>(target, acc, tid, a.owned, preload);
      run_init<R, NT, S>(acc, yb, tid, preload);
      target = yb;
    }
    lds_barrier();
    lds_stage<T, D, N, NT, EXACT, FWD, 0>(buf, a, b, tid);
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      int q = tid + NT * r;
      if (S % NT == 0 || q < S)
        acc[r] = acc[r] + buf[q];
    }
    keep_alive(pf);
  }
  run_flush<R, NT, S, 8>(target, acc, tid, a.owned, preload);
}

template<typename T, int D, int N, int NT, bool EXACT, bool FWD>
__global__ __launch_bounds__(NT) void kron_lds(EntryArgs<T> a)
{
  plan_bind(a);
  __shared__ T buf[cpow(N, D)];
  int p0, p1;
  wg_range(a, p0, p1);
  if (p0 >= p1)
    return;
  // owned runs with FMAs start from Y (see run_init), streamed X is loaded
  // non-temporally (gld_x); one specialised body per case keeps each loop's
  // code as tight as a kernel of its own.  EXACT keeps one plain body.
  if constexpr (EXACT)
    lds_body<T, D, N, NT, EXACT, FWD, false, false>(a, buf, p0, p1);
  else if (a.owned)
  {
    if (a.x_stream)
      lds_body<T, D, N, NT, EXACT, FWD, true, true>(a, buf, p0, p1);
    else
      lds_body<T, D, N, NT, EXACT, FWD, true, false>(a, buf, p0, p1);
  }
  else if (a.x_stream)
    lds_body<T, D, N, NT, EXACT, FWD, false, true>(a, buf, p0, p1);
  else
    lds_body<T, D, N, NT, EXACT, FWD, false, false>(a, buf, p0, p1);
}

// ---------------------------------------------------------------------------
// kron_reg: register-blocked passes (d6n4, d3n8, d4n4..d4n8).  The n^d tensor of
// an entry is held by NTH = n^(d-RM) threads with RR = n^RM elements each.
// Pass g keeps the modes [g*RM, (g+1)*RM) in registers and contracts them
// there (A_k in SGPRs, RR/n independent fibers per mode); between passes the
// tensor is transposed through a padded LDS buffer (one pad element every
// RR), so there are d/RM - 1 LDS round trips per entry instead of d.  Modes
// run 1..d (the recursive / forward order, kronmultv_recur.hpp:90-145).
// Layout g: thread t = (hi, lo) with hi the thread modes before the group and
// lo the L_g = n^(d-(g+1)RM) modes after it; register r = the group's modes;
// element index = hi*RR*L_g + r*L_g + lo.  Layout 0 makes every X load of a
// wavefront contiguous; outputs accumulate in the last layout and are
// transposed once per output run for a coalesced flush.
// ---------------------------------------------------------------------------
template<typename T, int D, int N, int RM, bool EXACT, int M>
__device__ __forceinline__ void reg_group(typename Acc<T>::type (&v)[cpow(N, RM)],
                                          const EntryArgs<T> &a, int b, int g)
{
  // contract group mode M (0 = slowest of the group): register stride N^(RM-1-M)
  using C = typename Acc<T>::type;
  constexpr int RR = cpow(N, RM);
  constexpr int ST = cpow(N, RM - 1 - M);
  constexpr int RB = (N <= 4) ? N : (N + 1) / 2; // A rows held in SGPRs at a time
  const T *Ak = a_of(a, b, g * RM + M, D);
  const int lda = a.lda ? a.lda : N;
  if constexpr (RB == N)
  {
    // all of A_k fits in SGPRs: contract each fiber in place
    C am[N][N];
    load_a<N, false, T>(am, Ak, lda);
#pragma unroll
    for (int base = 0; base < RR; base++)
    {
      if ((base / ST) % N != 0)
        continue;
      C x[N];
#pragma unroll
      for (int j = 0; j < N; j++)
        x[j] = v[base + j * ST];
#pragma unroll
      for (int i = 0; i < N; i++)
      {
        C acc;
        if constexpr (EXACT)
        {
          acc = 0;
#pragma unroll
          for (int j = 0; j < N; j++)
            acc = madd<true>(acc, am[i][j], x[j]);
        }
        else
        {
          acc = am[i][0] * x[0];
#pragma unroll
          for (int j = 1; j < N; j++)
            acc = madd<false>(acc, am[i][j], x[j]);
        }
        v[base + i * ST] = static_cast<C>(static_cast<T>(acc));
      }
    }
    if constexpr (M + 1 < RM)
      reg_group<T, D, N, RM, EXACT, M + 1>(v, a, b, g);
    return;
  }
  C o[RR];
#pragma unroll
  for (int r0 = 0; r0 < N; r0 += RB)
  {
    C am[RB][N];
#pragma unroll
    for (int j = 0; j < N; j++)
#pragma unroll
      for (int i = 0; i < RB; i++)
        if (r0 + i < N)
          am[i][j] = static_cast<C>(cld(Ak + (r0 + i + j * lda)));
#pragma unroll
    for (int base = 0; base < RR; base++)
    {
      if ((base / ST) % N != 0)
        continue; // fiber start: group digit M is zero
#pragma unroll
      for (int i = 0; i < RB; i++)
      {
        if (r0 + i >= N)
          continue; // odd n: the last block has fewer rows
        C acc;
        if constexpr (EXACT)
        {
          acc = 0;
#pragma unroll
          for (int j = 0; j < N; j++)
            acc = madd<true>(acc, am[i][j], v[base + j * ST]);
        }
        else
        {
          acc = am[i][0] * v[base];
#pragma unroll
          for (int j = 1; j < N; j++)
            acc = madd<false>(acc, am[i][j], v[base + j * ST]);
        }
        o[base + (r0 + i) * ST] = acc;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RR; r++)
    v[r] = static_cast<C>(static_cast<T>(o[r]));
  if constexpr (M + 1 < RM)
    reg_group<T, D, N, RM, EXACT, M + 1>(v, a, b, g);
}

template<int D, int N, int RM>
__device__ __forceinline__ int reg_index(int g, int t, int r)
{
  constexpr int RR = cpow(N, RM);
  // L_g = N^(D - (g+1)*RM), computed without a runtime power
  int L = 1;
#pragma unroll
  for (int k = 0; k < D; k++)
    if (k < D - (g + 1) * RM)
      L *= N;
  return (t / L) * RR * L + r * L + (t % L);
}

template<int RR>
__device__ __forceinline__ int reg_pad(int idx)
{
  return idx + idx / RR;
}

// L_g of reg_index as a compile-time value
template<int D, int N, int RM>
constexpr int reg_l(int g)
{
  int L = 1;
  for (int k = 0; k < D - (g + 1) * RM; k++)
    L *= N;
  return L;
}

// Padded position of register r of thread t in layout g: the position of
// register 0 plus r * reg_step(g).  With L_g a multiple of RR the pad term
// idx / RR grows by L_g / RR per register; with L_g = 1 the registers are
// idx = t*RR + r, r < RR, and the pad term stays t.  So each transpose needs
// one address per thread and compile-time offsets, instead of one address
// per register (which took VGPRs enough to spill at d4n8).
template<int D, int N, int RM>
constexpr int reg_step(int g)
{
  constexpr int RR = cpow(N, RM);
  const int L = reg_l<D, N, RM>(g);
  return L == 1 ? 1 : L + L / RR;
}

template<int D, int N, int RM>
constexpr bool reg_steps_exact()
{
  constexpr int RR = cpow(N, RM);
  for (int g = 0; g < D / RM; g++)
  {
    const int L = reg_l<D, N, RM>(g);
    if (L != 1 && L % RR != 0)
      return false;
  }
  return true;
}

template<typename T, int D, int N, int RM, bool EXACT, bool PRELOAD, bool XNT>
__device__ __forceinline__ void reg_body(const EntryArgs<T> &a, T *buf, int p0,
                                         int p1)
{
  using C = typename Acc<T>::type;
  constexpr int S = cpow(N, D);
  constexpr int RR = cpow(N, RM);
  constexpr int NTH = S / RR;
  constexpr int NG = D / RM;
  static_assert(D % RM == 0, "mode groups must tile d");
  static_assert(reg_steps_exact<D, N, RM>() && NTH % RR == 0,
                "transpose positions linear in the register index");
  // NTH need not be a whole number of wavefronts (n = 5, 6, 7): the last
  // wave runs with fewer lanes; every access below is indexed by tid < NTH
  const int tid = threadIdx.x;

  C v[RR];
  T acc[RR];
  T *target = nullptr;
  // an owned output run (a plan) starts from Y itself when FMAs are used, in
  // the accumulator's own layout (positions tid*RR + r), and ends with plain
  // stores; EXACT keeps Y + (sum) and shared outputs need atomics, so those
  // start from zero and leave through one coalesced LDS transpose
  constexpr bool preload = PRELOAD;

  auto flush = [&](T *dst) {
    if (preload)
    {
#pragma unroll
      for (int r = 0; r < RR; r++)
        gst(dst + tid * RR + r, acc[r]);
      return;
    }
    lds_barrier();
    {
      T *w = buf + reg_pad<RR>(reg_index<D, N, RM>(NG - 1, tid, 0));
#pragma unroll
      for (int r = 0; r < RR; r++)
        w[r * reg_step<D, N, RM>(NG - 1)] = acc[r];
    }
    lds_barrier();
    T out[RR];
    {
      // tid + NTH*k: NTH is a multiple of RR
      const T *q = buf + reg_pad<RR>(tid);
#pragma unroll
      for (int k = 0; k < RR; k++)
        out[k] = q[k * (NTH + NTH / RR)];
    }
    flush_run<RR, NTH, S, RR>(dst, out, tid, a.owned);
  };

  for (int p = p0; p < p1; ++p)
  {
    const int b = entry_at(a, p);
    // X first (layout 0: coalesced), so that waiting for it does not also
    // wait for the flush issued below
    const T *xp = x_of(a, b);
#pragma unroll
    for (int r = 0; r < RR; r++)
      v[r] = static_cast<C>(gld_x<XNT>(xp + r * NTH + tid));
    T *yb = y_of(a, b);
    if (yb != target)
    {
      if (target)
        flush(target);
#pragma unroll
      for (int r = 0; r < RR; r++)
        acc[r] = preload ? gld(yb + tid * RR + r) : T(0);
      target = yb;
    }
    // the next entry's A is pulled into L2 before mode group PF: with three
    // groups (d3n8) just before the last one, so that fewer of the lines
    // are evicted before the next entry's scalar loads (plan 0.494 -> 0.474
    // ms); with four or more, at the start (d4n8 measured slower later)
    constexpr int PF = (NG == 3) ? NG - 1 : 0;
    T pf = 0;
#pragma unroll
    for (int g = 0; g < NG; g++)
    {
      if (g == PF && kPrefetchA<N> && p + 1 < p1)
        pf = prefetch_a<D, N>(a, entry_at(a, p + 1), tid);
      reg_group<T, D, N, RM, EXACT, 0>(v, a, b, g);
      if (g + 1 < NG)
      {
        lds_barrier(); // previous readers of buf are done
        T *w = buf + reg_pad<RR>(reg_index<D, N, RM>(g, tid, 0));
#pragma unroll
        for (int r = 0; r < RR; r++)
          w[r * reg_step<D, N, RM>(g)] = static_cast<T>(v[r]);
        lds_barrier();
        const T *q = buf + reg_pad<RR>(reg_index<D, N, RM>(g + 1, tid, 0));
#pragma unroll
        for (int r = 0; r < RR; r++)
          v[r] = static_cast<C>(q[r * reg_step<D, N, RM>(g + 1)]);
      }
    }
#pragma unroll
    for (int r = 0; r < RR; r++)
      acc[r] = acc[r] + static_cast<T>(v[r]);
    keep_alive(pf);
  }
  flush(target);
}

template<typename T, int D, int N, int RM, bool EXACT>
__global__ __launch_bounds__(cpow(N, D - RM)) void kron_reg(EntryArgs<T> a)
{
  plan_bind(a);
  __shared__ T buf[cpow(N, D) + cpow(N, D - RM)];
  int p0, p1;
  wg_range(a, p0, p1);
  if (p0 >= p1)
    return;
  if constexpr (EXACT)
    reg_body<T, D, N, RM, EXACT, false, false>(a, buf, p0, p1);
  else if (a.owned)
  {
    if (a.x_stream)
      reg_body<T, D, N, RM, EXACT, true, true>(a, buf, p0, p1);
    else
      reg_body<T, D, N, RM, EXACT, true, false>(a, buf, p0, p1);
  }
  else if (a.x_stream)
    reg_body<T, D, N, RM, EXACT, false, true>(a, buf, p0, p1);
  else
    reg_body<T, D, N, RM, EXACT, false, false>(a, buf, p0, p1);
}

// ---------------------------------------------------------------------------
// kron_slab: X streamed in slabs X(j1, :) of n^(d-1) elements.
//   Y(i1, m) = sum_j1 A_1(i1, j1) * [kron(A_2..A_d) X(j1, :)](m)
// Inside a slab the modes d..2 are contracted in LDS (backward order); the
// mode-1 sum runs over j1 ascending straight into registers, so the order of
// every dot product is the Alg. 993 backward order.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void *lds_vptr;
typedef __attribute__((address_space(1))) void *gbl_vptr;

// Copy SL elements (SL*sizeof(T) a multiple of 16, src 16-byte aligned) from
// global memory straight into LDS, 16 bytes per lane.  Each wave writes one
// contiguous 1 KiB block per instruction.
template<typename T, int SL, int NT, bool XNT>
__device__ __forceinline__ void slab_async_load(T *dst, const T *src, int tid)
{
  if (reinterpret_cast<uintptr_t>(src) & 15)
  {
    // misaligned vector: ordinary element loads (still coalesced)
    for (int q = tid; q < SL; q += NT)
      dst[q] = gld(src + q);
    return;
  }
  constexpr int CH = SL * (int)sizeof(T) / 16; // 16-byte chunks
  constexpr int IT = (CH + NT - 1) / NT;
  const int wave = tid >> 6;
  const int lane = tid & 63;
#pragma unroll
  for (int it = 0; it < IT; it++)
  {
    int c0 = it * NT + wave * 64; // first chunk of this wave
    if (c0 < CH)
    {
      int c = c0 + lane;
      if (CH % 64 == 0 || c < CH)
      {
        const char *g = reinterpret_cast<const char *>(src) + (size_t)c * 16;
        char *l = reinterpret_cast<char *>(dst) + (size_t)c0 * 16;
        // cache-policy immediate: 2 = non-temporal
        __builtin_amdgcn_global_load_lds((gbl_vptr)(g), (lds_vptr)(l), 16, 0,
                                         XNT ? 2 : 0);
      }
    }
  }
}

template<typename T, int D, int N, int NT, bool EXACT, int K, bool FWD = false>
__device__ __forceinline__ void slab_stage(T *buf, const EntryArgs<T> &a, int b,
                                           int tid)
{
  // K runs 0..D-2 over the in-slab modes (global mode numbering): D-1 .. 1,
  // or 1 .. D-1 with FWD
  using C = typename Acc<T>::type;
  constexpr int SL = cpow(N, D - 1);
  constexpr int F = SL / N;
  constexpr int MODE = FWD ? 1 + K : D - 1 - K;
  constexpr int STRIDE = cpow(N, D - 1 - MODE);
  C am[N][N];
  // float: widened A back in SGPRs where n*n doubles fit there (n <= 6)
  load_a<N, (N <= 6)>(am, a_of(a, b, MODE, D), a.lda);
#pragma unroll kSlabUnroll
  for (int f = tid; f < F; f += NT)
  {
    int hi = f / STRIDE;
    int lo = f - hi * STRIDE;
    fiber_apply<N, EXACT, T>(buf + hi * STRIDE * N + lo, STRIDE, am);
  }
  lds_barrier();
  if constexpr (K + 2 < D)
    slab_stage<T, D, N, NT, EXACT, K + 1, FWD>(buf, a, b, tid);
}

// Pair layout of the slab accumulator: value r of a thread is element
// 2*(tid + NT*(r/2)) + r%2 of the slab, so each thread owns consecutive
// pairs and the run's Y preload and flush move 2 elements per lane and
// instruction (half the instructions against the per-wave limit of
// outstanding memory operations).  Used for owned runs that start from Y;
// shared runs keep one element per lane and instruction, which keeps each
// atomic instruction on contiguous addresses.
template<typename T>
using pair_t = T __attribute__((ext_vector_type(2)));

// Pair load/store of Y; an output vector that is not aligned to a pair
// (the caller's pointers need only element alignment) is accessed per
// element instead.
template<typename T>
__device__ __forceinline__ pair_t<T> pair_ld(const T *p, bool al)
{
  if (al)
    return *(gptr<const pair_t<T>>)(p);
  return pair_t<T>{gld(p), gld(p + 1)};
}
template<typename T>
__device__ __forceinline__ void pair_st(T *p, pair_t<T> v, bool al)
{
  if (al)
    *(gptr<pair_t<T>>)(p) = v;
  else
  {
    gst(p, v.x);
    gst(p + 1, v.y);
  }
}

template<int R, int NT, int S, typename T, typename C>
__device__ __forceinline__ void pair_init(C (&acc)[R], const T *src, int tid,
                                          bool preload)
{
  const bool al = reinterpret_cast<uintptr_t>(src) % sizeof(pair_t<T>) == 0;
#pragma unroll
  for (int rp = 0; rp < R / 2; rp++)
  {
    const int pr = tid + NT * rp;
    pair_t<T> v = {T(0), T(0)};
    if (preload && ((S / 2) % NT == 0 || pr < S / 2))
      v = pair_ld(src + 2 * pr, al);
    acc[2 * rp] = static_cast<C>(v.x);
    acc[2 * rp + 1] = static_cast<C>(v.y);
  }
}

template<int R, int NT, int S, typename T, typename C>
__device__ __forceinline__ void pair_flush(T *dst, const C (&acc)[R], int tid,
                                           int owned, bool preloaded)
{
  if (preloaded || owned)
  {
    const bool al = reinterpret_cast<uintptr_t>(dst) % sizeof(pair_t<T>) == 0;
    pair_t<T> old[R / 2];
    if (!preloaded)
    {
#pragma unroll
      for (int rp = 0; rp < R / 2; rp++)
      {
        const int pr = tid + NT * rp;
        old[rp] = {T(0), T(0)};
        if ((S / 2) % NT == 0 || pr < S / 2)
          old[rp] = pair_ld(dst + 2 * pr, al);
      }
    }
#pragma unroll
    for (int rp = 0; rp < R / 2; rp++)
    {
      const int pr = tid + NT * rp;
      if ((S / 2) % NT == 0 || pr < S / 2)
      {
        pair_t<T> v = {static_cast<T>(acc[2 * rp]), static_cast<T>(acc[2 * rp + 1])};
        if (!preloaded)
          v = old[rp] + v;
        pair_st(dst + 2 * pr, v, al);
      }
    }
  }
  else
  {
#pragma unroll
    for (int rp = 0; rp < R / 2; rp++)
    {
      const int pr = tid + NT * rp;
      if ((S / 2) % NT == 0 || pr < S / 2)
      {
        gatomic_add(dst + 2 * pr, static_cast<T>(acc[2 * rp]));
        gatomic_add(dst + 2 * pr + 1, static_cast<T>(acc[2 * rp + 1]));
      }
    }
  }
}

template<typename T, int D, int N, int NT, bool EXACT, bool PRELOAD, bool XNT>
__device__ __forceinline__ void slab_body(const EntryArgs<T> &a, T *sb0, T *sb1,
                                          int p0, int p1)
{
  constexpr int SL = cpow(N, D - 1);
  // measured: d6n6 plan 0.869 -> 0.771 ms; d5n8 plan 0.184 -> 0.186 ms
  constexpr bool PAIR = PRELOAD && N <= 6 && SL % 2 == 0;
  constexpr int RS = PAIR ? 2 * ((SL / 2 + NT - 1) / NT) : (SL + NT - 1) / NT;
  const int tid = threadIdx.x;

  using C = typename Acc<T>::type;
  C acc[N][RS];
  auto flush_all = [&](T *dst) {
#pragma unroll
    for (int i1 = 0; i1 < N; i1++)
    {
      if constexpr (PAIR)
        pair_flush<RS, NT, SL>(dst + i1 * SL, acc[i1], tid, a.owned, PRELOAD);
      else
        run_flush<RS, NT, SL, 2>(dst + i1 * SL, acc[i1], tid, a.owned, PRELOAD);
    }
  };
  T *target = nullptr;
  constexpr bool preload = PRELOAD;

  // slab sequence: (entry p, j1) flattened as t = (p - p0)*N + j1
  const int nslab = (p1 - p0) * N;
  // Workgroups run in near lockstep, so with equal output runs they would
  // all flush at the same moments, and HBM would see bursts of atomic Y
  // traffic between quiet stretches.  With shared (atomic) outputs each
  // workgroup therefore starts its range at position blockIdx % count and
  // wraps around: the run it starts inside is flushed twice, and the flushes
  // of neighbouring workgroups fall at different times.  Owned runs keep the
  // position order (a second flush of an owned run costs a Y read-modify-
  // write and measured slower).
  const int cnt = p1 - p0;
  const int rot = (PRELOAD || a.owned) ? 0 : (int)(blockIdx.x % (unsigned)cnt);
  auto pos = [&](int q) {
    if constexpr (!PRELOAD)
    {
      q += rot;
      q = q >= cnt ? q - cnt : q;
    }
    return p0 + q;
  };
  // Owned runs all switch at the same slab, and a switch (the old run's
  // stores plus the new run's Y loads, 2*N*SL elements) takes one CU about
  // 13 us beyond a plain slab even alone; with every CU at once it took
  // 23 us.  At d6n6 each workgroup therefore starts its slab sequence at
  // slab (5*blockIdx) mod nslab and wraps around: switches fall at
  // different times, and the run it starts inside is visited twice (its
  // second visit preloads what the first stored).  Measured: d6n6 plan
  // 0.770 -> 0.746 ms.  d5n8 has two runs per workgroup, so the extra visit
  // costs more than the spread saves (0.189 -> 0.194 ms): off there.
  constexpr int kStagger = (PRELOAD && D == 6) ? 5 : 0;
  const int srot =
      kStagger ? (int)((blockIdx.x * (unsigned)kStagger) % (unsigned)nslab) : 0;
  auto slab_of = [&](int t, int &j) {
    int u = t + srot;
    u = u >= nslab ? u - nslab : u;
    j = u - (u / N) * N;
    return pos(u / N);
  };
  {
    int j0;
    const int p0s = slab_of(0, j0);
    const T *xp = x_of(a, entry_at(a, p0s)) + (size_t)j0 * SL;
    slab_async_load<T, SL, NT, XNT>(sb0, xp, tid);
  }

  for (int t = 0; t < nslab; ++t)
  {
    int j1;
    const int p = slab_of(t, j1);
    const int b = entry_at(a, p);
    T *cur = (t & 1) ? sb1 : sb0;
    T *nxt = (t & 1) ? sb0 : sb1;

    // the current slab has landed for every wave; nxt is free again.  The
    // wait also covers the stores of the previous run's flush, which were
    // issued a whole slab earlier (below), so they drain behind compute.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (j1 == 0 || t == 0)
    {
      T *yb = y_of(a, b);
      if (yb != target)
      {
        if (target)
          flush_all(target);
#pragma unroll
        for (int i1 = 0; i1 < N; i1++)
        {
          if constexpr (PAIR)
            pair_init<RS, NT, SL>(acc[i1], yb + i1 * SL, tid, preload);
          else
            run_init<RS, NT, SL>(acc[i1], yb + i1 * SL, tid, preload);
        }
        target = yb;
      }
    }
    if (t + 1 < nslab)
    {
      int jn;
      const int pn = slab_of(t + 1, jn);
      const T *xp = x_of(a, entry_at(a, pn));
      slab_async_load<T, SL, NT, XNT>(nxt, xp + (size_t)jn * SL, tid);
    }

    slab_stage<T, D, N, NT, EXACT, 0>(cur, a, b, tid);

    const T *A1 = a_of(a, b, 0, D);
    C a1[N];
#pragma unroll
    for (int i1 = 0; i1 < N; i1++)
      a1[i1] = cld(A1 + (i1 + j1 * a.lda));
    if constexpr (PAIR)
    {
#pragma unroll
      for (int rp = 0; rp < RS / 2; rp++)
      {
        const int pr = tid + NT * rp;
        if ((SL / 2) % NT == 0 || pr < SL / 2)
        {
          const pair_t<T> zz = *reinterpret_cast<const pair_t<T> *>(cur + 2 * pr);
          const C z0 = zz.x, z1 = zz.y;
#pragma unroll
          for (int i1 = 0; i1 < N; i1++)
          {
            acc[i1][2 * rp] = madd<EXACT>(acc[i1][2 * rp], a1[i1], z0);
            acc[i1][2 * rp + 1] = madd<EXACT>(acc[i1][2 * rp + 1], a1[i1], z1);
          }
        }
      }
    }
    else
    {
#pragma unroll
      for (int r = 0; r < RS; r++)
      {
        int q = tid + NT * r;
        if (SL % NT == 0 || q < SL)
        {
          C z = cur[q];
#pragma unroll
          for (int i1 = 0; i1 < N; i1++)
            acc[i1][r] = madd<EXACT>(acc[i1][r], a1[i1], z);
        }
      }
    }
  }
  flush_all(target);
}

template<typename T, int D, int N, int NT, bool EXACT>
__global__ __launch_bounds__(NT) void kron_slab(EntryArgs<T> a)
{
  plan_bind(a);
  constexpr int SL = cpow(N, D - 1);
  static_assert((SL * sizeof(T)) % 16 == 0, "slab must be whole 16-byte chunks");
  __shared__ __attribute__((aligned(16))) T sb0[SL];
  __shared__ __attribute__((aligned(16))) T sb1[SL];
  int p0, p1;
  wg_range(a, p0, p1);
  if (p0 >= p1)
    return;
  // owned FMA runs start from Y; shared outputs of streamed X use
  // non-temporal slab loads (with owned runs those measured slower)
  if constexpr (EXACT)
    slab_body<T, D, N, NT, EXACT, false, false>(a, sb0, sb1, p0, p1);
  else if (a.owned)
    slab_body<T, D, N, NT, EXACT, true, false>(a, sb0, sb1, p0, p1);
  else if (a.x_stream)
    slab_body<T, D, N, NT, EXACT, false, true>(a, sb0, sb1, p0, p1);
  else
    slab_body<T, D, N, NT, EXACT, false, false>(a, sb0, sb1, p0, p1);
}

// ---------------------------------------------------------------------------
// kron_slab_fwd: the forward mode order 1..d with the reference's rounding
// (KM_FLAG_EXACT with KM_VARIANT_FORWARD, the reference's default recursive
// build).  Mode 1 comes first here, so the slabs X(j1, :) are streamed into
// a register accumulator of the mode-1 result Z(i1, :) = sum_j1 A_1(i1, j1)
// X(j1, :) (j1 ascending, separate multiply and add); then each output slab
// Z(i1, :) goes to LDS, modes 2..d are contracted there in order, and the
// slab is added to Y(i1, :).  Same passes and bytes as kron_slab; outputs
// are updated once per entry (no run merging: the registers hold Z).
// ---------------------------------------------------------------------------
template<typename T, int D, int N, int NT>
__device__ __forceinline__ void slab_fwd_body(const EntryArgs<T> &a, T *sb0,
                                              T *sb1, int p0, int p1)
{
  constexpr int SL = cpow(N, D - 1);
  constexpr int RS = (SL + NT - 1) / NT;
  const int tid = threadIdx.x;
  using C = typename Acc<T>::type;
  C acc[N][RS];

  const int nslab = (p1 - p0) * N;
  slab_async_load<T, SL, NT, false>(sb0, x_of(a, entry_at(a, p0)), tid);
  for (int t = 0; t < nslab; ++t)
  {
    const int p = p0 + t / N;
    const int j1 = t - (t / N) * N;
    const int b = entry_at(a, p);
    T *cur = (t & 1) ? sb1 : sb0;
    T *nxt = (t & 1) ? sb0 : sb1;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (t + 1 < nslab)
    {
      const int pn = p0 + (t + 1) / N;
      const int jn = (t + 1) - ((t + 1) / N) * N;
      slab_async_load<T, SL, NT, false>(nxt, x_of(a, entry_at(a, pn)) + (size_t)jn * SL,
                                        tid);
    }
    const T *A1 = a_of(a, b, 0, D);
    C a1[N];
#pragma unroll
    for (int i1 = 0; i1 < N; i1++)
      a1[i1] = cld(A1 + (i1 + j1 * a.lda));
#pragma unroll
    for (int r = 0; r < RS; r++)
    {
      const int q = tid + NT * r;
      if (SL % NT == 0 || q < SL)
      {
        const C x = cur[q];
#pragma unroll
        for (int i1 = 0; i1 < N; i1++)
          acc[i1][r] = madd<true>(j1 == 0 ? C(0) : acc[i1][r], a1[i1], x);
      }
    }
    if (j1 == N - 1)
    {
      // the entry's output slabs, one at a time, in the buffer of its last
      // input slab (the next DMA goes to the other buffer)
      T *yb = y_of(a, b);
#pragma unroll
      for (int i1 = 0; i1 < N; i1++)
      {
        lds_barrier(); // readers of cur are done
#pragma unroll
        for (int r = 0; r < RS; r++)
        {
          const int q = tid + NT * r;
          if (SL % NT == 0 || q < SL)
            cur[q] = static_cast<T>(acc[i1][r]);
        }
        lds_barrier();
        slab_stage<T, D, N, NT, true, 0, true>(cur, a, b, tid);
#pragma unroll
        for (int r = 0; r < RS; r++)
        {
          const int q = tid + NT * r;
          if (SL % NT == 0 || q < SL)
          {
            T *yp = yb + (size_t)i1 * SL + q;
            if (a.owned)
            {
              // the previous entry of this output was written by this thread;
              // a device-coherent load does not see a stale cached line
              const T y = __hip_atomic_load((gptr<T>)(yp), __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
              gst(yp, y + cur[q]);
            }
            else
              gatomic_add(yp, cur[q]);
          }
        }
      }
    }
  }
}

template<typename T, int D, int N, int NT>
__global__ __launch_bounds__(NT) void kron_slab_fwd(EntryArgs<T> a)
{
  plan_bind(a);
  constexpr int SL = cpow(N, D - 1);
  static_assert((SL * sizeof(T)) % 16 == 0, "slab must be whole 16-byte chunks");
  __shared__ __attribute__((aligned(16))) T sb0[SL];
  __shared__ __attribute__((aligned(16))) T sb1[SL];
  int p0, p1;
  wg_range(a, p0, p1);
  if (p0 >= p1)
    return;
  slab_fwd_body<T, D, N, NT>(a, sb0, sb1, p0, p1);
}

// ---------------------------------------------------------------------------
// Two-pass family, for shapes whose n^(d-1) slab exceeds LDS (d = 6, n = 7, 8):
//   pass 1 (kron_sub):  for every entry b and every sub-slab (j1, j2) the
//     trailing modes d..3 are contracted in LDS and the result
//     Z_b(j1, j2, :) = kron(A_3..A_d) X_b(j1, j2, :) is stored to the entry's
//     workspace (the caller's W, as in the reference);
//   pass 2 (kron_pair): per position m of the trailing modes, the n x n values
//     Z_b(:, :, m) are contracted over mode 2 and then mode 1 in registers and
//     added into Y.
// The mode order is d, d-1, ..., 1: the Alg. 993 backward order.
// ---------------------------------------------------------------------------
template<typename T>
__device__ __forceinline__ T *w_of(const EntryArgs<T> &a, int b, int local)
{
  const int s = a.slot_by_entry ? b : local;
  return a.pW ? cld(a.pW + a.ws_mul * s) : a.W + (long long)s * a.w_stride;
}

// One mode pass over F fibers of stride STRIDE in LDS: every thread first
// loads all of its fibers (independent LDS reads in flight), then applies A
// in blocks of RB rows so that at most RB*N entries of A sit in SGPRs (an
// 8 x 8 FP64 matrix alone would need 128), then writes the fibers back.
template<typename T, int N, int NT, bool EXACT, int F, int STRIDE>
__device__ __forceinline__ void stage_apply(T *buf, const T *Ak, int lda,
                                            int tid)
{
  using C = typename Acc<T>::type;
  constexpr int FPT = (F + NT - 1) / NT;
  constexpr int RB = (N <= 6) ? N : (N + 1) / 2;
  C x[FPT][N];
  C o[FPT][N];
#pragma unroll
  for (int f = 0; f < FPT; f++)
  {
    const int fi = tid + NT * f;
    if (F % NT == 0 || fi < F)
    {
      const int hi = fi / STRIDE, lo = fi - (fi / STRIDE) * STRIDE;
      const T *p = buf + hi * STRIDE * N + lo;
#pragma unroll
      for (int j = 0; j < N; j++)
        x[f][j] = static_cast<C>(p[j * STRIDE]);
    }
  }
#pragma unroll
  for (int r0 = 0; r0 < N; r0 += RB)
  {
    C am[RB][N];
#pragma unroll
    for (int j = 0; j < N; j++)
#pragma unroll
      for (int i = 0; i < RB; i++)
        if (r0 + i < N)
          am[i][j] = static_cast<C>(cld(Ak + (r0 + i + j * lda)));
#pragma unroll
    for (int f = 0; f < FPT; f++)
#pragma unroll
      for (int i = 0; i < RB; i++)
        if (r0 + i < N)
        {
          C acc = 0;
          if constexpr (EXACT)
          {
#pragma unroll
            for (int j = 0; j < N; j++)
              acc = madd<true>(acc, am[i][j], x[f][j]);
          }
          else
          {
            acc = am[i][0] * x[f][0];
#pragma unroll
            for (int j = 1; j < N; j++)
              acc = madd<false>(acc, am[i][j], x[f][j]);
          }
          o[f][r0 + i] = acc;
        }
  }
#pragma unroll
  for (int f = 0; f < FPT; f++)
  {
    const int fi = tid + NT * f;
    if (F % NT == 0 || fi < F)
    {
      const int hi = fi / STRIDE, lo = fi - (fi / STRIDE) * STRIDE;
      T *p = buf + hi * STRIDE * N + lo;
#pragma unroll
      for (int i = 0; i < N; i++)
        p[i * STRIDE] = static_cast<T>(o[f][i]);
    }
  }
}

// Contract modes MODE, MODE-1, ..., LAST of an LDS tensor that holds modes
// LAST..D-1 of one entry.
template<typename T, int D, int N, int NT, bool EXACT, int LAST, int MODE>
__device__ __forceinline__ void trailing_stages(T *buf, const EntryArgs<T> &a,
                                                int b, int tid)
{
  constexpr int SZ = cpow(N, D - LAST);
  constexpr int F = SZ / N;
  constexpr int STRIDE = cpow(N, D - 1 - MODE);
  stage_apply<T, N, NT, EXACT, F, STRIDE>(buf, a_of(a, b, MODE, D), a.lda, tid);
  lds_barrier();
  if constexpr (MODE > LAST)
    trailing_stages<T, D, N, NT, EXACT, LAST, MODE - 1>(buf, a, b, tid);
}

// pass 1: sub-entries s in [0, nsub) of the chunk starting at order position
// e0; sub-entry s is sub-slab s % (N*N) of the entry at position e0 + s/(N*N)
template<typename T, int D, int N, int NT, bool EXACT>
__global__ __launch_bounds__(NT) void kron_sub(EntryArgs<T> a, int e0, int nsub)
{
  constexpr int M = cpow(N, D - 2);
  constexpr int R = (M + NT - 1) / NT;
  constexpr int NN = N * N;
  __shared__ T buf[M];
  const int tid = threadIdx.x;
  const int s0 = blockIdx.x * a.chunk;
  const int s1 = min(nsub, s0 + a.chunk);
  if (s0 >= s1)
    return;
  T xr[R];
  {
    const int b = entry_at(a, e0 + s0 / NN);
    const T *xp = x_of(a, b) + (long long)(s0 % NN) * M;
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      int q = tid + NT * r;
      if (M % NT == 0 || q < M)
        xr[r] = gld(xp + q);
    }
  }
  for (int s = s0; s < s1; ++s)
  {
    const int p = e0 + s / NN;
    const int js = s % NN;
    const int b = entry_at(a, p);
    lds_barrier();
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      int q = tid + NT * r;
      if (M % NT == 0 || q < M)
        buf[q] = xr[r];
    }
    if (s + 1 < s1)
    {
      const int bn = entry_at(a, e0 + (s + 1) / NN);
      const T *xp = x_of(a, bn) + (long long)((s + 1) % NN) * M;
#pragma unroll
      for (int r = 0; r < R; r++)
      {
        int q = tid + NT * r;
        if (M % NT == 0 || q < M)
          xr[r] = gld(xp + q);
      }
    }
    lds_barrier();
    trailing_stages<T, D, N, NT, EXACT, 2, D - 1>(buf, a, b, tid);
    T *wb = w_of(a, b, p - e0) + (long long)js * M;
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      int q = tid + NT * r;
      if (M % NT == 0 || q < M)
        gst(wb + q, buf[q]);
    }
  }
}

// pass 2: blockIdx.y selects NT consecutive positions m; blockIdx.x selects a
// range of order positions (plan ranges, or chunks of a.chunk from e0).
template<typename T, int D, int N, int NT, bool EXACT>
__global__ __launch_bounds__(NT) void kron_pair(EntryArgs<T> a, int e0, int e1)
{
  plan_bind(a);
  using C = typename Acc<T>::type;
  constexpr int M = cpow(N, D - 2);
  const int tid = threadIdx.x;
  const int m = blockIdx.y * NT + tid;
  int p0, p1;
  if (a.ranges)
  {
    p0 = cld(a.ranges + blockIdx.x);
    p1 = cld(a.ranges + blockIdx.x + 1);
  }
  else
  {
    p0 = e0 + blockIdx.x * a.chunk;
    p1 = min(e1, p0 + a.chunk);
  }
  const bool live = (M % NT == 0) || (m < M);
  // Mode 1 is applied as rank-1 updates, one row j1 of W at a time:
  //   acc(i1, :) += A_1(i1, j1) * [A_2 W(j1, :)^T]^T,
  // so only the n x n accumulator and two n-value rows are live (each acc
  // element still sums its terms over j1 ascending).  FMA calls sum runs of
  // consecutive entries with the same output in registers and leave with
  // one store (owned, started from Y) or one atomic per element per run;
  // EXACT keeps the reference's Y + (entry sum), one update per entry.
  C acc[N][N];
  T *target = nullptr;
  auto flush = [&](T *yb) {
    if (!live)
      return;
#pragma unroll
    for (int i1 = 0; i1 < N; i1++)
#pragma unroll
      for (int i2 = 0; i2 < N; i2++)
      {
        T *yp = yb + (long long)(i1 * N + i2) * M + m;
        const T v = static_cast<T>(acc[i1][i2]);
        if (!a.owned)
          gatomic_add(yp, v);
        else if constexpr (!EXACT)
          gst(yp, v);
        else
        {
          // the previous entry of this output was written by this thread;
          // a device-coherent load does not see a stale cached line
          const T y = __hip_atomic_load((gptr<T>)(yp), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
          gst(yp, y + v);
        }
      }
  };
  for (int p = p0; p < p1; ++p)
  {
    const int b = entry_at(a, p);
    T *yb = y_of(a, b);
    if (EXACT || yb != target)
    {
      if (target)
        flush(target);
      // an owned FMA run starts from Y itself
#pragma unroll
      for (int i1 = 0; i1 < N; i1++)
#pragma unroll
        for (int i2 = 0; i2 < N; i2++)
          acc[i1][i2] = (!EXACT && a.owned && live)
                            ? static_cast<C>(gld(yb + (long long)(i1 * N + i2) * M + m))
                            : C(0);
      target = yb;
    }
    const T *wb = w_of(a, b, p - e0);
    const T *A1 = a_of(a, b, 0, D);
    const T *A2 = a_of(a, b, 1, D);
#pragma unroll 1
    for (int j1 = 0; j1 < N; j1++)
    {
      C w[N];
#pragma unroll
      for (int j2 = 0; j2 < N; j2++)
        w[j2] = live ? static_cast<C>(gld(wb + (long long)(j1 * N + j2) * M + m)) : C(0);
      C u[N];
#pragma unroll
      for (int i2 = 0; i2 < N; i2++)
      {
        C t = 0;
#pragma unroll
        for (int j2 = 0; j2 < N; j2++)
          t = madd<EXACT>(t, static_cast<C>(cld(A2 + (i2 + j2 * a.lda))), w[j2]);
        u[i2] = static_cast<C>(static_cast<T>(t));
      }
#pragma unroll
      for (int i1 = 0; i1 < N; i1++)
      {
        const C a1 = static_cast<C>(cld(A1 + (i1 + j1 * a.lda)));
#pragma unroll
        for (int i2 = 0; i2 < N; i2++)
          acc[i1][i2] = madd<EXACT>(acc[i1][i2], a1, u[i2]);
      }
    }
  }
  if (target)
    flush(target);
}

// ---------------------------------------------------------------------------
// Forward two-pass family (KM_FLAG_EXACT with KM_VARIANT_FORWARD, where one
// slab overflows LDS): the mirror of kron_sub + kron_pair for the mode order
// 1..d.
//   pass 1 (kron_pair_fwd): per position m of the trailing modes, modes 1
//     then 2 of the n x n values X(:, :, m) in registers, each result
//     rounded to T as a stage output, into W(i1, i2, m);
//   pass 2 (kron_sub_fwd): per (i1, i2) sub-slab W(i1, i2, :) in LDS, modes
//     3..d in order, then added to Y(i1, i2, :) -- atomically, or with a plan
//     as a read-modify-write per entry in entry order (blockIdx.y fixes the
//     sub-slab, so one workgroup updates each part of an owned output).
// ---------------------------------------------------------------------------
template<typename T, int D, int N, int NT, int MODE>
__device__ __forceinline__ void forward_stages(T *buf, const EntryArgs<T> &a,
                                               int b, int tid)
{
  constexpr int SZ = cpow(N, D - 2);
  constexpr int F = SZ / N;
  constexpr int STRIDE = cpow(N, D - 1 - MODE);
  stage_apply<T, N, NT, true, F, STRIDE>(buf, a_of(a, b, MODE, D), a.lda, tid);
  lds_barrier();
  if constexpr (MODE + 1 < D)
    forward_stages<T, D, N, NT, MODE + 1>(buf, a, b, tid);
}

template<typename T, int D, int N, int NT>
__global__ __launch_bounds__(NT) void kron_pair_fwd(EntryArgs<T> a, int e0, int e1)
{
  using C = typename Acc<T>::type;
  constexpr int M = cpow(N, D - 2);
  const int tid = threadIdx.x;
  const int m = blockIdx.y * NT + tid;
  const int p0 = e0 + blockIdx.x * a.chunk;
  const int p1 = min(e1, p0 + a.chunk);
  if (!((M % NT == 0) || (m < M)))
    return;
  for (int p = p0; p < p1; ++p)
  {
    const int b = entry_at(a, p);
    const T *xb = x_of(a, b);
    T *wb = w_of(a, b, p - e0);
    C x[N][N];
#pragma unroll
    for (int j1 = 0; j1 < N; j1++)
#pragma unroll
      for (int j2 = 0; j2 < N; j2++)
        x[j1][j2] = static_cast<C>(gld(xb + (long long)(j1 * N + j2) * M + m));
    // mode 1: x(i1, j2) <- sum_j1 A_1(i1, j1) x(j1, j2), one column at a time
    const T *A1 = a_of(a, b, 0, D);
#pragma unroll
    for (int j2 = 0; j2 < N; j2++)
    {
      C t[N];
#pragma unroll
      for (int i1 = 0; i1 < N; i1++)
      {
        C acc = 0;
#pragma unroll
        for (int j1 = 0; j1 < N; j1++)
          acc = madd<true>(acc, static_cast<C>(cld(A1 + (i1 + j1 * a.lda))), x[j1][j2]);
        t[i1] = static_cast<C>(static_cast<T>(acc));
      }
#pragma unroll
      for (int i1 = 0; i1 < N; i1++)
        x[i1][j2] = t[i1];
    }
    // mode 2: W(i1, i2) = sum_j2 A_2(i2, j2) x(i1, j2), one row at a time
    const T *A2 = a_of(a, b, 1, D);
#pragma unroll
    for (int i1 = 0; i1 < N; i1++)
    {
#pragma unroll
      for (int i2 = 0; i2 < N; i2++)
      {
        C acc = 0;
#pragma unroll
        for (int j2 = 0; j2 < N; j2++)
          acc = madd<true>(acc, static_cast<C>(cld(A2 + (i2 + j2 * a.lda))), x[i1][j2]);
        gst(wb + (long long)(i1 * N + i2) * M + m, static_cast<T>(acc));
      }
    }
  }
}

template<typename T, int D, int N, int NT>
__global__ __launch_bounds__(NT) void kron_sub_fwd(EntryArgs<T> a, int e0, int e1)
{
  plan_bind(a);
  constexpr int M = cpow(N, D - 2);
  constexpr int R = (M + NT - 1) / NT;
  __shared__ T buf[M];
  const int tid = threadIdx.x;
  const int js = blockIdx.y; // sub-slab (i1, i2) = (js / N, js % N)
  int p0, p1;
  if (a.ranges)
  {
    p0 = cld(a.ranges + blockIdx.x);
    p1 = cld(a.ranges + blockIdx.x + 1);
  }
  else
  {
    p0 = e0 + blockIdx.x * a.chunk;
    p1 = min(e1, p0 + a.chunk);
  }
  if (p0 >= p1)
    return;
  T wr[R];
  {
    const T *wp = w_of(a, entry_at(a, p0), p0 - e0) + (long long)js * M;
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      const int q = tid + NT * r;
      if (M % NT == 0 || q < M)
        wr[r] = gld(wp + q);
    }
  }
  for (int p = p0; p < p1; ++p)
  {
    const int b = entry_at(a, p);
    lds_barrier(); // buf free
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      const int q = tid + NT * r;
      if (M % NT == 0 || q < M)
        buf[q] = wr[r];
    }
    if (p + 1 < p1)
    {
      const T *wp = w_of(a, entry_at(a, p + 1), p + 1 - e0) + (long long)js * M;
#pragma unroll
      for (int r = 0; r < R; r++)
      {
        const int q = tid + NT * r;
        if (M % NT == 0 || q < M)
          wr[r] = gld(wp + q);
      }
    }
    lds_barrier();
    forward_stages<T, D, N, NT, 2>(buf, a, b, tid);
    T *yb = y_of(a, b) + (long long)js * M;
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      const int q = tid + NT * r;
      if (M % NT == 0 || q < M)
      {
        if (a.owned)
        {
          // the previous entry of this output was written by this thread;
          // a device-coherent load does not see a stale cached line
          const T y = __hip_atomic_load((gptr<T>)(yb + q), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
          gst(yb + q, y + buf[q]);
        }
        else
          gatomic_add(yb + q, buf[q]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// kron_global: any d <= 6 and any (rectangular) m_k x n_k; one workgroup per
// entry; stages go through the caller's workspace as in the reference
// (kronm_backward.hpp:92-139).  Square shapes update one buffer in place;
// rectangular shapes ping-pong between the slot's two buffers.  The final
// stage adds into Y with atomics, or with a plain read-modify-write when the
// workgroup owns the output (the barrier after every stage orders the
// entries of one output inside the workgroup).
// ---------------------------------------------------------------------------
template<typename T, bool EXACT>
__global__ __launch_bounds__(256) void kron_global(EntryArgs<T> a, Shape sh,
                                                   int fwd)
{
  plan_bind(a);
  using C = typename Acc<T>::type;
  extern __shared__ __attribute__((aligned(16))) unsigned char kg_dyn[];
  const int tid = threadIdx.x;
  const int d = sh.d;
  int nmax = 0;
  for (int k = 0; k < d; k++)
    nmax = max(nmax, sh.n[k]);
  int p0, p1;
  wg_range(a, p0, p1);
  bool square = true;
  for (int k = 0; k < d; k++)
    square = square && (sh.m[k] == sh.n[k]);
  T *w0 = a.pW ? cld(a.pW + a.ws_mul * blockIdx.x)
               : a.W + blockIdx.x * a.w_stride;
  T *w1 = square ? w0
                 : (a.pW ? cld(a.pW + (a.ws_mul * blockIdx.x + 1))
                         : a.W + blockIdx.x * a.w_stride + a.w_half);

  for (int p = p0; p < p1; ++p)
  {
    const int b = entry_at(a, p);
    const T *src = x_of(a, b);
    T *yb = y_of(a, b);
    // current tensor extents: mode k has extent ext[k]
    int ext[6];
    for (int k = 0; k < d; k++)
      ext[k] = sh.n[k];
    T *bufs[2] = {w0, w1};
    int which = 0;
    for (int s = 0; s < d; s++)
    {
      const int k = fwd ? s : d - 1 - s;
      const int nk = sh.n[k], mk = sh.m[k];
      int inner = 1;
      for (int q = k + 1; q < d; q++)
        inner *= ext[q];
      int outer = 1;
      for (int q = 0; q < k; q++)
        outer *= ext[q];
      const T *Ak = a_of(a, b, k, d);
      const int ld = a.lda ? a.lda : mk;
      const bool last = (s == d - 1);
      T *dst = bufs[which];
      const long long fibers = (long long)outer * inner;
      for (long long f = tid; f < fibers; f += blockDim.x)
      {
        long long hi = f / inner, lo = f - hi * inner;
        const T *xin = src + hi * inner * nk + lo;
        if (nk <= KM_GLOBAL_MAX_N)
        {
          // the fiber in registers
          C x[KM_GLOBAL_MAX_N];
          for (int j = 0; j < nk; j++)
            x[j] = static_cast<C>(gld(xin + (long long)j * inner));
          for (int i = 0; i < mk; i++)
          {
            C acc = 0;
            for (int j = 0; j < nk; j++)
              acc = madd<EXACT>(acc, static_cast<C>(cld(Ak + (i + j * ld))), x[j]);
            long long o = hi * inner * mk + lo + (long long)i * inner;
            if (last)
              flush_add(yb + o, static_cast<T>(acc), a.owned);
            else
              gst(dst + o, static_cast<T>(acc));
          }
        }
        else
        {
          // wider modes: the fiber goes to this thread's slice of the
          // dynamic LDS buffer (blockDim.x * nmax values), so square shapes
          // may still update their buffer in place
          C *xf = reinterpret_cast<C *>(kg_dyn) + (size_t)tid * nmax;
          for (int j = 0; j < nk; j++)
            xf[j] = static_cast<C>(gld(xin + (long long)j * inner));
          for (int i = 0; i < mk; i++)
          {
            C acc = 0;
            for (int j = 0; j < nk; j++)
              acc = madd<EXACT>(acc, static_cast<C>(cld(Ak + (i + j * ld))), xf[j]);
            long long o = hi * inner * mk + lo + (long long)i * inner;
            if (last)
              flush_add(yb + o, static_cast<T>(acc), a.owned);
            else
              gst(dst + o, static_cast<T>(acc));
          }
        }
      }
      __syncthreads();
      ext[k] = mk;
      src = dst;
      if (!square)
        which ^= 1;
    }
  }
}

} // namespace km
