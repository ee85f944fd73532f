// km_runs.hip -- kron_runs: the grouped variable-size batch (per-entry n) in
// one launch.  Each workgroup owns whole output runs (all entries of one
// output, consecutive in the group plan's order: sizes descending, then
// grouped by output), so every output is updated by one workgroup with a
// plain read-modify-write, bitwise reproducible from call to call.
//
// Ways through a run, chosen by the entry size S = n^d:
//   large (S above one wavefront's LDS region): the whole workgroup
//       contracts one entry at a time, the run's sum kept in registers.
//       - double-buffered: contracted in LDS (one fiber per thread and
//         mode); LDS holds two entries: while one is contracted, the next
//         entry of the workgroup's range is copied in by direct-to-LDS loads
//         and its factors are pulled into L2, so neither waits on global
//         memory and no registers hold them.
//       - register-blocked (FMA calls at d = 4, the largest size n = 8):
//         kron_reg's mode groups, X loaded straight into registers, one LDS
//         transpose between modes (runs_reg_unit).
//   small: the run's entries are dealt out to the wavefronts (entry k of the
//       run to wave k mod waves); each wave loads its entry into its own LDS
//       region and contracts it with wave-level synchronisation only, summing
//       its entries in registers, while its next entry's X and factors are
//       pulled into L2.  At the end of the run the waves' partial sums meet
//       in LDS and are added to Y in a fixed order.
// An LDS mode pass loads every fiber of a thread first and applies the
// factor a few rows at a time from scalar registers (stage_apply), so n = 7,
// 8 need no more SGPRs than one block of rows.  The entries of a run and the
// waves' partial sums are added in a fixed order, so the result does not
// depend on timing.  Mode order: d..1, the Alg. 993 backward order
// (kronm_backward.hpp:92-139), as kron_lds; the register-blocked body runs
// 1..d as kron_reg (FMA calls only, where the order changes rounding alone).
// The kernel also serves the small sizes of owned per-size grouped calls
// (the plan's ranges over those runs alone).
#include "km_internal.h"

namespace km
{

// Whether large entries take the register-blocked body (runs_reg_unit): FMA
// calls at d = 4 with 512 threads, where the large sizes n = 6, 7, 8 have
// n^3 = 216, 343, 512 <= 512 threads of n values each.
template<int D, int NT, bool EXACT>
constexpr bool runs_reg_geom()
{
  return !EXACT && D == 4 && NT == 512;
}

constexpr int cmax(int x, int y) { return x > y ? x : y; }

// REG: entries of the largest size n = NMAX take the register-blocked body,
// whose LDS is one padded transpose buffer (n^d + n^(d-1)); the other large
// ones are double-buffered in LDS as below, in two halves of the same
// buffer; small entries get 636-element regions (n <= 5 at d = 4): 40 KiB in
// all.  Otherwise every large entry is double-buffered in LDS (2 n^d).
template<typename T, int D, int NMAX, int NT, bool REG = false>
struct RunsGeom
{
  static constexpr int SMAX = cpow(NMAX, D);               // largest entry
  static constexpr int NW = NT / 64;                        // wavefronts
  static constexpr int SWR = 636;                           // REG: per-wave region
  static constexpr int LDSE = REG ? cmax(SMAX + cpow(NMAX, D - 1), NW * SWR)
                                  : 2 * SMAX;               // LDS elements
  static constexpr int SW = REG ? SWR : (NW > 1 ? (LDSE / NW) & ~3 : 0);  // per-wave region
  static constexpr int DMAX = REG ? cpow(NMAX - 1, D) : SMAX; // largest double-buffered entry
  static constexpr int BSTR = REG ? (LDSE / 2) & ~1 : SMAX; // offset of the second buffer
  static_assert(DMAX <= BSTR, "two entries fit");
  static constexpr int RB = (DMAX + NT - 1) / NT;           // large-entry sum per thread
  static constexpr int RS = NW > 1 ? (SW + 63) / 64 : 1;    // small-entry sum per lane
};

__device__ __forceinline__ int ipow_dev(int n, int d)
{
  int s = 1;
  for (int k = 0; k < d; k++)
    s *= n;
  return s;
}

// Copy nelem elements at src into LDS at dst with direct-to-LDS loads:
// 16 bytes per lane when src and dst are 16-byte aligned and the bytes are
// whole 16-byte chunks, else 4 bytes.  Threads [0, NTH) take part (tid
// relative to them: the workgroup or one wavefront); each wavefront writes
// one contiguous block of LDS per instruction.  Completion is tracked by
// vmcnt.
template<typename T, int NTH>
__device__ __forceinline__ void copy_to_lds(T *dst, const T *src, int nelem, int tid)
{
  const int bytes = nelem * (int)sizeof(T);
  const int wave0 = (tid & ~63); // first thread of this wavefront
  const int lane = tid & 63;
  if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0 &&
      (bytes & 15) == 0)
  {
    const int ch = bytes >> 4;
    for (int c0 = wave0; c0 < ch; c0 += NTH)
    {
      const int c = c0 + lane;
      if (c < ch)
        __builtin_amdgcn_global_load_lds(
            (gbl_vptr)(reinterpret_cast<const char *>(src) + (size_t)c * 16),
            (lds_vptr)(reinterpret_cast<char *>(dst) + (size_t)c0 * 16), 16, 0, 0);
    }
  }
  else
  {
    const int ch = bytes >> 2;
    for (int c0 = wave0; c0 < ch; c0 += NTH)
    {
      const int c = c0 + lane;
      if (c < ch)
        __builtin_amdgcn_global_load_lds(
            (gbl_vptr)(reinterpret_cast<const char *>(src) + (size_t)c * 4),
            (lds_vptr)(reinterpret_cast<char *>(dst) + (size_t)c0 * 4), 4, 0, 0);
    }
  }
}

// Pull lines into L2 without registers: direct-to-LDS loads (4 bytes per
// lane, one lane per 128-byte line) into a scratch block of LDS that nothing
// reads.  Lanes [0, 64) of one wavefront; elems elements at p.
template<typename T>
__device__ __forceinline__ void pull_lines(const T *p, int elems, int lane, int *scratch)
{
  constexpr int PL = 128 / (int)sizeof(T);
  const int lines = (elems + PL - 1) / PL;
  for (int l0 = 0; l0 < lines; l0 += 64)
    if (l0 + lane < lines)
      __builtin_amdgcn_global_load_lds(
          (gbl_vptr)(p + (size_t)(l0 + lane) * PL), (lds_vptr)(scratch), 4, 0, 0);
}

// the d factors of entry b (n x n each, stored densely) into L2
template<typename T, int D>
__device__ __forceinline__ void pull_a(const EntryArgs<T> &a, int b, int n, int lane,
                                       int *scratch)
{
#pragma unroll
  for (int k = 0; k < D; k++)
    pull_lines(a_of(a, b, k, D), n * n, lane, scratch);
}

// Y(0:S) += acc, thread tid holding elements tid + NTH*r: plain
// read-modify-write for an owned output, atomics otherwise.
template<int R, int NTH, typename T>
__device__ __forceinline__ void runs_flush(T *dst, const T (&acc)[R], int tid,
                                           int S, int owned)
{
  if (owned)
  {
    T yv[R];
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      const int q = tid + NTH * r;
      if (q < S)
        yv[r] = gld(dst + q);
    }
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      const int q = tid + NTH * r;
      if (q < S)
        gst(dst + q, yv[r] + acc[r]);
    }
  }
  else
  {
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      const int q = tid + NTH * r;
      if (q < S)
        gatomic_add(dst + q, acc[r]);
    }
  }
}

// modes d..1 of the entry in buf (N^D elements, LDS) with NTH threads
template<typename T, int D, int N, int NTH, bool EXACT, int K, bool WAVE>
__device__ __forceinline__ void runs_stages(T *buf, const EntryArgs<T> &a, int b, int tid)
{
  constexpr int MODE = D - 1 - K;
  constexpr int STRIDE = cpow(N, D - 1 - MODE);
  constexpr int F = cpow(N, D) / N;
  stage_apply<T, N, NTH, EXACT, F, STRIDE>(buf, a_of(a, b, MODE, D), N, tid);
  pass_sync<WAVE>();
  if constexpr (K + 1 < D)
    runs_stages<T, D, N, NTH, EXACT, K + 1, WAVE>(buf, a, b, tid);
}

// contract the entry in buf (size N^D) with NTH threads and add it to acc
template<typename T, int D, int N, int NTH, int R, bool EXACT, bool WAVE>
__device__ __forceinline__ void runs_entry(T *buf, const EntryArgs<T> &a, int b, int tid,
                                           T (&acc)[R])
{
  constexpr int S = cpow(N, D);
  runs_stages<T, D, N, NTH, EXACT, 0, WAVE>(buf, a, b, tid);
#pragma unroll
  for (int r = 0; r < (S + NTH - 1) / NTH; r++)
  {
    const int q = tid + NTH * r;
    if (S % NTH == 0 || q < S)
      acc[r] = acc[r] + buf[q];
  }
}

// dispatch an entry of size n to its compile-time body: small bodies (the
// entry fits one wavefront's region) with one wavefront, large ones with the
// workgroup
template<typename T, int D, int NMAX, int NT, bool EXACT, bool LARGE, int R>
__device__ __forceinline__ void runs_dispatch(int n, T *buf, const EntryArgs<T> &a, int b,
                                              int tid, T (&acc)[R])
{
  using G = RunsGeom<T, D, NMAX, NT, runs_reg_geom<D, NT, EXACT>()>;
  constexpr int NTH = LARGE ? NT : 64;
  switch (n)
  {
#define KM_RUNS_N(NN)                                                          \
  case NN:                                                                     \
    if constexpr (NN <= NMAX && (LARGE ? ((G::NW == 1 || cpow(NN, D) > G::SW) && \
                                          cpow(NN, D) <= G::DMAX)                 \
                                       : (G::NW > 1 && cpow(NN, D) <= G::SW)))    \
      runs_entry<T, D, NN, NTH, R, EXACT, !LARGE>(buf, a, b, tid, acc);        \
    break;
    KM_RUNS_N(1)
    KM_RUNS_N(2)
    KM_RUNS_N(3)
    KM_RUNS_N(4)
    KM_RUNS_N(5)
    KM_RUNS_N(6)
    KM_RUNS_N(7)
    KM_RUNS_N(8)
#undef KM_RUNS_N
  default:
    break;
  }
}

// Part sum of a split run: stored to the plan's workspace; kron_runs_combine,
// launched right after on the same stream, adds the parts to Y in part order.
template<int R, int NT, typename T>
__device__ __forceinline__ void runs_part_store(const RunsArgs<T> &ra, int sid, int part,
                                                const T (&acc)[R], int tid, int S)
{
  T *mine = ra.ws + cld(ra.wsoff + sid) + (long long)part * S;
#pragma unroll
  for (int r = 0; r < R; r++)
  {
    const int q = tid + NT * r;
    if (q < S)
      gst(mine + q, acc[r]);
  }
}

// One workgroup per split run: Y += part 0 + part 1 + ... (in that order),
// so the result does not depend on which part finished first.
template<typename T, int D, int NT>
__global__ __launch_bounds__(NT) void kron_runs_combine(EntryArgs<T> a, RunsArgs<T> ra)
{
  plan_bind(a);
  if (!a.owned)
    return; // the caller's outputs differ from the plan's: parts went to Y with atomics
  const int sid = blockIdx.x;
  const int b = entry_at(a, cld(ra.sfirst + sid));
  const int S = ipow_dev(cld(ra.nvec + b), D);
  const int parts = cld(ra.sparts + sid);
  const T *base = ra.ws + cld(ra.wsoff + sid);
  T *y = y_of(a, b);
  for (int q = threadIdx.x; q < S; q += NT)
  {
    T v = gld(base + q);
    for (int k = 1; k < parts; k++)
      v = v + gld(base + (long long)k * S + q);
    gst(y + q, gld(y + q) + v);
  }
}


// One work unit [ustart[cu], ustart[cu+1]) of size-N entries with the
// register-blocked body: thread t < N^(D-1) loads X(r*N^(D-1) + t), contracts
// modes 1..d in registers with one LDS transpose between modes (buf, padded
// as in kron_reg), and sums the unit's entries in registers in the last
// layout, where it holds elements t*N .. t*N + N-1.  The other threads
// contract zeros and read and write a spare LDS slot instead: inside the
// entry loop no branch depends on the thread, so the factors stay in scalar
// registers (a lane-divergent branch there made the compiler load them per
// lane, 61 -> 180 VGPRs).  Forward mode order, as kron_reg (FMA calls only:
// the order changes rounding alone).  OWNED: the plan's outputs (a plain
// read-modify-write, or a part sum of a split run, at the end of the unit);
// otherwise every change of output is flushed with atomics.
template<typename T, int D, int N, int NT, bool OWNED>
__device__ __forceinline__ void runs_reg_unit(const EntryArgs<T> &a, const RunsArgs<T> &ra,
                                              int cu, T *buf, int tid_in)
{
  // tid through an empty asm statement: the per-thread offsets below are
  // computed per unit, not hoisted out of the caller's loop over units
  // (where the offsets of all three sizes stayed live: 116 VGPRs)
  int tid = tid_in;
  asm volatile("" : "+v"(tid));
  using C = typename Acc<T>::type;
  constexpr int RM = 1;
  constexpr int RR = N;
  constexpr int S = cpow(N, D);
  constexpr int NTH = cpow(N, D - 1);
  constexpr int NG = D;
  // padded transpose positions end below S + NTH; the spare slot follows
  constexpr int SPARE = S + NTH;
  static_assert(NTH <= NT && SPARE + 1 <= RunsGeom<T, D, 8, NT, true>::LDSE,
                "the unit fits the workgroup and its LDS");
  const bool act = (NTH == NT) || tid < NTH;
  const int t = act ? tid : 0; // idle threads read X(r*NTH) (discarded)
  const int p0 = cld(ra.ustart + cu), p1 = cld(ra.ustart + cu + 1);
  T acc[RR];
#pragma unroll
  for (int r = 0; r < RR; r++)
    acc[r] = 0;
  T *target = y_of(a, entry_at(a, p0));
  for (int p = p0; p < p1; ++p)
  {
    const int b = entry_at(a, p);
    if constexpr (!OWNED)
    {
      // the caller's outputs differ from the plan's: flush where they change
      T *yb = y_of(a, b);
      if (yb != target)
      {
        if (act)
        {
#pragma unroll
          for (int r = 0; r < RR; r++)
            gatomic_add(target + tid * RR + r, acc[r]);
        }
#pragma unroll
        for (int r = 0; r < RR; r++)
          acc[r] = 0;
        target = yb;
      }
    }
    C v[RR];
    {
      const T *xp = x_of(a, b);
#pragma unroll
      for (int r = 0; r < RR; r++)
        v[r] = static_cast<C>(gld(xp + r * NTH + t));
    }
    T pf = 0;
#pragma unroll
    for (int g = 0; g < NG; g++)
    {
      if (g == 0 && kPrefetchA<N> && p + 1 < p1)
        pf = prefetch_a<D, N>(a, entry_at(a, p + 1), tid);
      reg_group<T, D, N, RM, false, 0>(v, a, b, g);
      if (g + 1 < NG)
      {
        lds_barrier(); // previous readers of buf are done
        {
          T *w = buf + (act ? reg_pad<RR>(reg_index<D, N, RM>(g, tid, 0)) : SPARE);
          const int ST = reg_step<D, N, RM>(g);
#pragma unroll
          for (int r = 0; r < RR; r++)
            w[act ? r * ST : 0] = static_cast<T>(v[r]);
        }
        lds_barrier();
        {
          const T *q = buf + (act ? reg_pad<RR>(reg_index<D, N, RM>(g + 1, tid, 0)) : SPARE);
          const int ST = reg_step<D, N, RM>(g + 1);
#pragma unroll
          for (int r = 0; r < RR; r++)
            v[r] = static_cast<C>(q[act ? r * ST : 0]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RR; r++)
      acc[r] = acc[r] + static_cast<T>(v[r]);
    keep_alive(pf);
  }
  if (!act)
    return;
  if constexpr (!OWNED)
  {
#pragma unroll
    for (int r = 0; r < RR; r++)
      gatomic_add(target + tid * RR + r, acc[r]);
    return;
  }
  const int sid = cld(ra.umeta + 3 * cu);
  if (sid < 0)
  {
    // owned output: read-modify-write of this thread's N elements
    T yv[RR];
#pragma unroll
    for (int r = 0; r < RR; r++)
      yv[r] = gld(target + tid * RR + r);
#pragma unroll
    for (int r = 0; r < RR; r++)
      gst(target + tid * RR + r, yv[r] + acc[r]);
  }
  else
  {
    // one part of a split run: its sum to the plan's workspace
    T *mine = ra.ws + cld(ra.wsoff + sid) + (long long)cld(ra.umeta + 3 * cu + 1) * S;
#pragma unroll
    for (int r = 0; r < RR; r++)
      gst(mine + tid * RR + r, acc[r]);
  }
}

// With the register-blocked large entries, the kernel is compiled for three
// workgroups per CU in FP64 (at most 80 VGPRs; two values of the
// double-buffered sum go to 32 bytes of scratch), which its 40 KiB of LDS
// allows (unconstrained it took 96 VGPRs: two per CU), and for two in float
// (at most 128 VGPRs; unconstrained 138: one per CU; three per CU spilled
// 34).
template<typename T, int D, int NT, bool EXACT>
constexpr int runs_waves()
{
  return runs_reg_geom<D, NT, EXACT>() ? (sizeof(T) == 8 ? 6 : 4) : 1;
}

template<typename T, int D, int NMAX, int NT, bool EXACT>
__global__ __launch_bounds__(NT)
__attribute__((amdgpu_waves_per_eu(runs_waves<T, D, NT, EXACT>()))) void kron_runs(EntryArgs<T> a, RunsArgs<T> ra)
{
  constexpr bool REG = runs_reg_geom<D, NT, EXACT>();
  using G = RunsGeom<T, D, NMAX, NT, REG>;
  plan_bind(a);
  __shared__ __attribute__((aligned(16))) T lds[G::LDSE];
  __shared__ int scratch[64];
  const int tid = threadIdx.x;
  int u0, u1;
  wg_range(a, u0, u1);
  if (u0 >= u1)
    return;
  const int *nvec = ra.nvec;
  auto small = [&](int n) { return G::NW > 1 && ipow_dev(n, D) <= G::SW; };

  // units of large entries come first (sizes descend along the plan's order)
  int ub = u0;
  while (ub < u1 && !small(cld(nvec + entry_at(a, cld(ra.ustart + ub)))))
    ++ub;

  // FMA calls at d = 4: the units of the largest size (n = 8, first in
  // the plan's order, where sizes descend) with the register-blocked body,
  // one unit at a time
  int ur = u0;
  if constexpr (REG)
  {
    auto unit_n = [&](int u) { return cld(nvec + entry_at(a, cld(ra.ustart + u))); };
    if (a.owned)
    {
      for (; ur < ub && unit_n(ur) == NMAX; ++ur)
        runs_reg_unit<T, D, NMAX, NT, true>(a, ra, ur, lds, tid);
    }
    else
    {
      for (; ur < ub && unit_n(ur) == NMAX; ++ur)
        runs_reg_unit<T, D, NMAX, NT, false>(a, ra, ur, lds, tid);
    }
    if (ur > u0 && ur < ub)
      lds_barrier(); // the last transposes are read before the buffers are reused
  }

  if (ub > ur)
  {
    const int q0 = cld(ra.ustart + ur), qb = cld(ra.ustart + ub);
    T acc[G::RB];
#pragma unroll
    for (int r = 0; r < G::RB; r++)
      acc[r] = 0;
    int b = entry_at(a, q0);
    int n = cld(nvec + b);
    // the unit (or, when the caller's outputs differ from the plan's, the
    // output run) being summed
    T *target = y_of(a, b);
    int tn = n, tsize = ipow_dev(n, D);
    int cu = ur, ce = cld(ra.ustart + ur + 1);
    auto flush = [&]() {
      if (!a.owned)
      {
        runs_flush<G::RB, NT>(target, acc, tid, tsize, 0);
        return;
      }
      const int sid = cld(ra.umeta + 3 * cu);
      if (sid < 0)
        runs_flush<G::RB, NT>(target, acc, tid, tsize, 1);
      else
        runs_part_store<G::RB, NT>(ra, sid, cld(ra.umeta + 3 * cu + 1), acc, tid, tsize);
    };
    copy_to_lds<T, NT>(lds, x_of(a, b), tsize, tid);
    if (tid < 64)
      pull_a<T, D>(a, b, n, tid, scratch);
    for (int p = q0; p < qb; ++p)
    {
      // the next entry's pointers, loaded while this entry waits
      int bn = b, nn = n;
      const T *xn = nullptr;
      if (p + 1 < qb)
      {
        bn = entry_at(a, p + 1);
        nn = cld(nvec + bn);
        xn = x_of(a, bn);
      }
      // offsets, not a select of two pointers: a select lets the compiler
      // turn every LDS read into reads of both buffers and a select
      const int par = (p - q0) & 1;
      T *cur = lds + par * G::BSTR;
      T *nxt = lds + (par ^ 1) * G::BSTR;
      if (p > q0)
      {
        T *yb = y_of(a, b);
        const bool unit_end = p == ce;
        if (a.owned ? unit_end : (yb != target || n != tn))
        {
          flush();
#pragma unroll
          for (int r = 0; r < G::RB; r++)
            acc[r] = 0;
        }
        if (unit_end)
        {
          ++cu;
          ce = cld(ra.ustart + cu + 1);
        }
        target = yb;
        tn = n;
        tsize = ipow_dev(n, D);
      }
      // this thread's copies into cur have landed; after the barrier all
      // have, and every thread is done with nxt (the previous entry)
      __builtin_amdgcn_s_waitcnt(0);
      lds_barrier();
      if (p + 1 < qb)
      {
        // the next entry: X into the other buffer, its factors into L2
        copy_to_lds<T, NT>(nxt, xn, ipow_dev(nn, D), tid);
        if (tid < 64)
          pull_a<T, D>(a, bn, nn, tid, scratch);
      }
      runs_dispatch<T, D, NMAX, NT, EXACT, true>(n, cur, a, b, tid, acc);
      b = bn;
      n = nn;
    }
    flush();
  }

  if constexpr (G::NW > 1)
  {
    // the wavefront index is uniform over the wavefront: made explicit, so
    // everything derived from it (entries, pointers, factors) stays scalar
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane_in = tid & 63;
    T *wbuf = lds + wave * G::SW;
    for (int r = ub; r < u1; r++)
    {
      // one output run [p, re): entries of one output and one size
      const int p = cld(ra.ustart + r), re = cld(ra.ustart + r + 1);
      const int b0 = entry_at(a, p);
      T *yb = y_of(a, b0);
      const int n = cld(nvec + b0);
      const int S = ipow_dev(n, D);
      // the regions are free: every wave is done with the previous run
      lds_barrier();
      // outputs not owned (the per-size path's atomic calls, or a caller's
      // pY that differs from the plan's): the run is still summed on chip
      // when all its entries have the same output, and flushed with atomics
      bool same = a.owned;
      if (!a.owned)
      {
        int diff = 0;
        for (int e2 = p + wave; e2 < re; e2 += G::NW)
          diff |= (y_of(a, entry_at(a, e2)) != yb) ? 1 : 0;
        same = !__syncthreads_or(diff);
      }
      T acc[G::RS];
#pragma unroll
      for (int q = 0; q < G::RS; q++)
        acc[q] = 0;
      int e = p + wave;
      int b = e < re ? entry_at(a, e) : 0;
      const T *xp = e < re ? x_of(a, b) : nullptr;
      for (; e < re; e += G::NW)
      {
        // the lane through an empty asm statement: the per-lane offsets of
        // every size's passes are computed per entry, not hoisted out of the
        // loops (where they were spilled to scratch and reloaded per entry)
        int lane = lane_in;
        asm volatile("" : "+v"(lane));
        copy_to_lds<T, 64>(wbuf, xp, S, lane);
        // this wave's next entry of the run: pointers, its X and factors
        // into L2
        int bn = b;
        const T *xn = nullptr;
        if (e + G::NW < re)
        {
          bn = entry_at(a, e + G::NW);
          xn = x_of(a, bn);
          pull_lines(xn, S, lane, scratch);
          pull_a<T, D>(a, bn, n, lane, scratch);
        }
        __builtin_amdgcn_s_waitcnt(0);
        pass_sync<true>();
        runs_dispatch<T, D, NMAX, NT, EXACT, false>(n, wbuf, a, b, lane, acc);
        pass_sync<true>(); // the region is read before it is reloaded
        if (!same)
        {
          // the run's entries have different outputs: each entry goes to
          // its own output with atomics
          runs_flush<G::RS, 64>(y_of(a, b), acc, lane, S, 0);
#pragma unroll
          for (int q = 0; q < G::RS; q++)
            acc[q] = 0;
        }
        b = bn;
        xp = xn;
      }
      if (!same)
        continue;
      const int used = min(G::NW, re - p);
      if (wave < used)
      {
#pragma unroll
        for (int q = 0; q < G::RS; q++)
        {
          const int i = lane_in + 64 * q;
          if (i < S)
            wbuf[i] = acc[q];
        }
      }
      lds_barrier();
      // fixed order: wave 0's partial first, then 1, ..
      constexpr int RT = (G::SW + NT - 1) / NT;
      T sum[RT];
#pragma unroll
      for (int q = 0; q < RT; q++)
      {
        const int i = tid + NT * q;
        sum[q] = 0;
        if (i < S)
        {
          T v = lds[i];
          for (int w = 1; w < used; w++)
            v = v + lds[w * G::SW + i];
          sum[q] = v;
        }
      }
      runs_flush<RT, NT>(yb, sum, tid, S, a.owned);
    }
  }
}

#define KM_RUNS_CASE(T, D, NMAX, NT)                                           \
  if (d == D)                                                                  \
  {                                                                            \
    out->fn = exact ? (const void *)&kron_runs<T, D, NMAX, NT, true>           \
                    : (const void *)&kron_runs<T, D, NMAX, NT, false>;         \
    out->nt = NT;                                                              \
    out->family = FAM_GROUPED;                                                 \
    *nmax = NMAX;                                                              \
    *small_max = exact ? RunsGeom<T, D, NMAX, NT>::SW                          \
                       : RunsGeom<T, D, NMAX, NT, runs_reg_geom<D, NT, false>()>::SW; \
    return true;                                                               \
  }

template<>
const void *runs_combine_kernel<double>(int d, int *nt)
{
  *nt = 256;
  switch (d)
  {
  case 1: return (const void *)&kron_runs_combine<double, 1, 256>;
  case 2: return (const void *)&kron_runs_combine<double, 2, 256>;
  case 3: return (const void *)&kron_runs_combine<double, 3, 256>;
  case 4: return (const void *)&kron_runs_combine<double, 4, 256>;
  case 5: return (const void *)&kron_runs_combine<double, 5, 256>;
  case 6: return (const void *)&kron_runs_combine<double, 6, 256>;
  default: return nullptr;
  }
}

template<>
const void *runs_combine_kernel<float>(int d, int *nt)
{
  *nt = 256;
  switch (d)
  {
  case 1: return (const void *)&kron_runs_combine<float, 1, 256>;
  case 2: return (const void *)&kron_runs_combine<float, 2, 256>;
  case 3: return (const void *)&kron_runs_combine<float, 3, 256>;
  case 4: return (const void *)&kron_runs_combine<float, 4, 256>;
  case 5: return (const void *)&kron_runs_combine<float, 5, 256>;
  case 6: return (const void *)&kron_runs_combine<float, 6, 256>;
  default: return nullptr;
  }
}

template<>
bool runs_kernel<double>(int d, bool exact, KernelInfo *out, int *nmax, int *small_max)
{
  KM_RUNS_CASE(double, 1, 8, 64)
  KM_RUNS_CASE(double, 2, 8, 64)
  KM_RUNS_CASE(double, 3, 8, 256)
  KM_RUNS_CASE(double, 4, 8, 512)
  KM_RUNS_CASE(double, 5, 6, 512)
  KM_RUNS_CASE(double, 6, 4, 512)
  return false;
}

template<>
bool runs_kernel<float>(int d, bool exact, KernelInfo *out, int *nmax, int *small_max)
{
  KM_RUNS_CASE(float, 1, 8, 64)
  KM_RUNS_CASE(float, 2, 8, 64)
  KM_RUNS_CASE(float, 3, 8, 256)
  KM_RUNS_CASE(float, 4, 8, 512)
  KM_RUNS_CASE(float, 5, 6, 512)
  KM_RUNS_CASE(float, 6, 4, 512)
  return false;
}

} // namespace km
