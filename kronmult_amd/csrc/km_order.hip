// km_order.hip -- output-ordered entry order built on the device, and the
// plan binding check.
//
// The order: a stable LSD radix sort of the entries by their output pointer
// (six passes of 8-bit digits over address bits 0..47, entry index as the
// value, so entries of one output keep ascending entry order), then the run
// starts of the sorted keys and the cut of the order into at most G
// workgroup ranges that end only where an output's run ends (see ord_cuts).
// Every buffer is written before it is read, so nothing needs clearing.  No
// counterpart in the reference (its overlapping outputs use atomics,
// kgemm_nt.hpp:184-187).
#include <algorithm>

#include "km_internal.h"

namespace km
{
namespace
{

typedef unsigned long long u64;

constexpr int kOT = 256;             // threads per workgroup
constexpr int kOPer = 16;            // entries per thread
constexpr int kOBlock = kOT * kOPer; // entries per workgroup
constexpr int kOPasses = 6;          // 8-bit digits over address bits 0..47

__global__ __launch_bounds__(kOT) void ord_init(const u64 *py, int batch,
                                                u64 *key, int *val)
{
  for (long long i = (long long)blockIdx.x * kOT + threadIdx.x; i < batch;
       i += (long long)gridDim.x * kOT)
  {
    key[i] = py[i];
    val[i] = (int)i;
  }
}

// per-workgroup digit counts: hist[digit * nblk + block]
__global__ __launch_bounds__(kOT) void ord_hist(const u64 *key, int batch,
                                                int shift, int nblk, int *hist)
{
  __shared__ int cnt[256];
  const int t = threadIdx.x;
  cnt[t] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kOBlock;
#pragma unroll 4
  for (int j = 0; j < kOPer; j++)
  {
    const long long i = base + j * kOT + t;
    if (i < batch)
      atomicAdd(&cnt[(int)((key[i] >> shift) & 255)], 1);
  }
  __syncthreads();
  hist[t * nblk + blockIdx.x] = cnt[t];
}

// Stable scatter of one workgroup's entries.  Entries are taken 256 at a
// time in index order; inside a wavefront the lanes holding the same digit
// are found with eight ballots, and a lane's rank is the number of such lanes
// below it; wavefront counts per digit go through LDS.
__global__ __launch_bounds__(kOT) void ord_scatter(const u64 *kin, const int *vin,
                                                   u64 *kout, int *vout, int batch,
                                                   int shift, int nblk,
                                                   const int *hist)
{
  __shared__ int base[256];
  __shared__ int scan[256];
  __shared__ int wcnt[kOT / 64][256];
  const int t = threadIdx.x;
  int tot = 0, pre = 0;
  for (int b = 0; b < nblk; b++)
  {
    const int h = hist[t * nblk + b];
    tot += h;
    pre += (b < (int)blockIdx.x) ? h : 0;
  }
  scan[t] = tot;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1)
  {
    const int v = (t >= off) ? scan[t - off] : 0;
    __syncthreads();
    scan[t] += v;
    __syncthreads();
  }
  base[t] = scan[t] - tot + pre;
#pragma unroll
  for (int w = 0; w < kOT / 64; w++)
    wcnt[w][t] = 0;
  __syncthreads();
  const int lane = t & 63, wave = t >> 6;
  const u64 below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const long long b0 = (long long)blockIdx.x * kOBlock;
  for (int j = 0; j < kOPer; j++)
  {
    const long long i = b0 + j * kOT + t;
    const bool valid = i < batch;
    const u64 k = valid ? kin[i] : 0ull;
    const int dg = (int)((k >> shift) & 255);
    u64 m = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; bit++)
    {
      const u64 bb = __ballot(valid && ((dg >> bit) & 1));
      m &= ((dg >> bit) & 1) ? bb : ~bb;
    }
    const int rank = __popcll(m & below);
    if (valid && rank == 0)
      wcnt[wave][dg] = __popcll(m);
    __syncthreads();
    if (valid)
    {
      int off = base[dg] + rank;
      for (int w = 0; w < wave; w++)
        off += wcnt[w][dg];
      kout[off] = k;
      vout[off] = vin[i];
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int w = 0; w < kOT / 64; w++)
    {
      add += wcnt[w][t];
      wcnt[w][t] = 0;
    }
    base[t] += add;
    __syncthreads();
  }
}

// rs[i] = last run start at or before i inside the workgroup's part (-1 if
// none); per-workgroup last run start and number of runs
__global__ __launch_bounds__(kOT) void ord_runs(const u64 *key, int batch, int *rs,
                                                int *blkmax, int *blkcnt)
{
  __shared__ int smax[kOT];
  __shared__ int ssum[kOT];
  const int t = threadIdx.x;
  const long long base = (long long)blockIdx.x * kOBlock + (long long)t * kOPer;
  int last = -1, cnt = 0;
  for (int q = 0; q < kOPer; q++)
  {
    const long long i = base + q;
    if (i < batch)
    {
      if (i == 0 || key[i] != key[i - 1])
      {
        last = (int)i;
        cnt++;
      }
      rs[i] = last;
    }
  }
  smax[t] = last;
  ssum[t] = cnt;
  __syncthreads();
  for (int off = 1; off < kOT; off <<= 1)
  {
    const int vm = (t >= off) ? smax[t - off] : -1;
    const int vs = (t >= off) ? ssum[t - off] : 0;
    __syncthreads();
    smax[t] = max(smax[t], vm);
    ssum[t] += vs;
    __syncthreads();
  }
  const int before = t ? smax[t - 1] : -1;
  for (int q = 0; q < kOPer; q++)
  {
    const long long i = base + q;
    if (i < batch && rs[i] < 0)
      rs[i] = before;
  }
  if (t == kOT - 1)
  {
    blkmax[blockIdx.x] = smax[t];
    blkcnt[blockIdx.x] = ssum[t];
  }
}

// Workgroup ranges.  Cut k (0 < k < G) is the start of the run holding
// position k*batch/G; cuts that fall inside one run coincide.  The distinct
// cuts are packed to the front, so the D - 1 non-empty ranges go to
// consecutive workgroups (which the dispatcher deals out round-robin over
// the XCDs) and the remaining workgroups get empty ranges [batch, batch).
// out[0] = number of outputs, out[1] = number of non-empty ranges.
__global__ __launch_bounds__(kOT) void ord_cuts(const int *rs, const int *blkmax,
                                                const int *blkcnt, int nblk,
                                                int batch, int G, int *ranges,
                                                int *out)
{
  __shared__ int cut[kOrderMaxGroups + 1];
  __shared__ int sc[kOT];
  const int t = threadIdx.x;
  for (int k = t; k <= G; k += kOT)
  {
    int r;
    if (k == 0)
      r = 0;
    else if (k == G)
      r = batch;
    else
    {
      const long long ideal = (long long)k * batch / G;
      r = rs[ideal];
      if (r < 0)
      {
        const int blk = (int)(ideal / kOBlock);
        for (int b = 0; b < blk; b++)
          r = max(r, blkmax[b]);
      }
    }
    cut[k] = r;
  }
  __syncthreads();
  // thread t packs the cuts of its segment [k0, k1)
  const int per = (G + 1 + kOT - 1) / kOT;
  const int k0 = min(G + 1, t * per), k1 = min(G + 1, k0 + per);
  int c = 0;
  for (int k = k0; k < k1; k++)
    c += (k == 0 || cut[k] != cut[k - 1]) ? 1 : 0;
  sc[t] = c;
  __syncthreads();
  for (int off = 1; off < kOT; off <<= 1)
  {
    const int v = (t >= off) ? sc[t - off] : 0;
    __syncthreads();
    sc[t] += v;
    __syncthreads();
  }
  int j = sc[t] - c;
  for (int k = k0; k < k1; k++)
    if (k == 0 || cut[k] != cut[k - 1])
      ranges[j++] = cut[k];
  const int D = sc[kOT - 1];
  for (int k = D + t; k <= G; k += kOT)
    ranges[k] = batch;
  if (t == 0)
  {
    int s = 0;
    for (int b = 0; b < nblk; b++)
      s += blkcnt[b];
    out[0] = s;
    out[1] = D - 1;
  }
}

// flag = gen when any pY[b] differs from the pointer the plan recorded
__global__ __launch_bounds__(kOT) void plan_check(const u64 *py, const u64 *ybind,
                                                  int batch, int *flag,
                                                  unsigned gen)
{
  bool bad = false;
  for (long long i = (long long)blockIdx.x * kOT + threadIdx.x; i < batch;
       i += (long long)gridDim.x * kOT)
    bad = bad || (py[i] != ybind[i]);
  if (bad)
    *flag = (int)gen;
}

struct Layout
{
  size_t key[2], val[2], rs, hist, blkmax, blkcnt, ranges, nout, total;
};

size_t up(size_t v) { return (v + 255) & ~(size_t)255; }

Layout layout(long long batch, int G)
{
  const long long nblk = (batch + kOBlock - 1) / kOBlock;
  Layout L;
  size_t o = 0;
  for (int i = 0; i < 2; i++)
  {
    L.key[i] = o;
    o = up(o + sizeof(u64) * batch);
  }
  for (int i = 0; i < 2; i++)
  {
    L.val[i] = o;
    o = up(o + sizeof(int) * batch);
  }
  L.rs = o;
  o = up(o + sizeof(int) * batch);
  L.hist = o;
  o = up(o + sizeof(int) * 256 * nblk);
  L.blkmax = o;
  o = up(o + sizeof(int) * nblk);
  L.blkcnt = o;
  o = up(o + sizeof(int) * nblk);
  L.ranges = o;
  o = up(o + sizeof(int) * (G + 1));
  L.nout = o;
  o = up(o + 2 * sizeof(int));
  L.total = o;
  return L;
}

} // namespace

size_t order_bytes(long long batch, int G) { return layout(batch, G).total; }

hipError_t order_build(const void *const *pY, int batch, int G, void *buf,
                       hipStream_t st, OrderOut *out)
{
  const int nblk = (batch + kOBlock - 1) / kOBlock;
  if (batch < 1 || G < 1 || G > kOrderMaxGroups || nblk > kOrderMaxBlocks)
    return hipErrorInvalidValue;
  const Layout L = layout(batch, G);
  char *base = static_cast<char *>(buf);
  u64 *key[2] = {reinterpret_cast<u64 *>(base + L.key[0]),
                 reinterpret_cast<u64 *>(base + L.key[1])};
  int *val[2] = {reinterpret_cast<int *>(base + L.val[0]),
                 reinterpret_cast<int *>(base + L.val[1])};
  int *rs = reinterpret_cast<int *>(base + L.rs);
  int *hist = reinterpret_cast<int *>(base + L.hist);
  int *blkmax = reinterpret_cast<int *>(base + L.blkmax);
  int *blkcnt = reinterpret_cast<int *>(base + L.blkcnt);
  int *ranges = reinterpret_cast<int *>(base + L.ranges);
  int *nout = reinterpret_cast<int *>(base + L.nout);

  const int ginit = std::min(1024, (batch + kOT - 1) / kOT);
  ord_init<<<ginit, kOT, 0, st>>>(reinterpret_cast<const u64 *>(pY), batch,
                                  key[0], val[0]);
  for (int p = 0; p < kOPasses; p++)
  {
    const int s = p & 1;
    ord_hist<<<nblk, kOT, 0, st>>>(key[s], batch, 8 * p, nblk, hist);
    ord_scatter<<<nblk, kOT, 0, st>>>(key[s], val[s], key[s ^ 1], val[s ^ 1],
                                      batch, 8 * p, nblk, hist);
  }
  static_assert(kOPasses % 2 == 0, "the sorted order ends in buffer 0");
  ord_runs<<<nblk, kOT, 0, st>>>(key[0], batch, rs, blkmax, blkcnt);
  ord_cuts<<<1, kOT, 0, st>>>(rs, blkmax, blkcnt, nblk, batch, G, ranges, nout);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return e;
  out->order = val[0];
  out->ranges = ranges;
  out->nout = nout;
  return hipSuccess;
}

hipError_t plan_check_launch(const void *const *pY, const void *const *ybind,
                             int batch, int *flag, unsigned gen, hipStream_t st)
{
  const int g = std::max(1, std::min(1024, (batch + kOT - 1) / kOT));
  plan_check<<<g, kOT, 0, st>>>(reinterpret_cast<const u64 *>(pY),
                                reinterpret_cast<const u64 *>(ybind), batch,
                                flag, gen);
  return hipGetLastError();
}

} // namespace km
