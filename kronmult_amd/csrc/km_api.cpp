// km_api.cpp -- the C ABI (include/kronmult_amd.h): host-side validation,
// kernel selection, launch geometry and deterministic-reduction plans.
//
// Entry points replace the reference's host wrappers:
//   km_kronmult_batched_*   kronmult6_batched.cpp:4-18   (and N = 1..5)
//   km_kronmult_pbatched_*  kronmult6_pbatched.cpp:4-19
//   km_kronmult_xbatched_*  kronmult6_xbatched.cpp:4-28
//   km_kronmult_vbatched_*  kronmult6_vbatched.hpp:10-34 (template only there)
#include <algorithm>
#include <atomic>
#include <link.h>
#include <sys/stat.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <utility>
#include <vector>

#include "../../include/kronmult_amd.h"
#include "km_internal.h"

using km::EntryArgs;
using km::KernelInfo;

struct km_plan_s
{
  int d = 0, n = 0, elem_bytes = 0, batch = 0;
  int device = 0;
  int ngroups = 0;
  int noutputs = 0;
  int owned = 1; // 1: output-ordered (owned outputs); 0: input-ordered (atomics)
  int *order = nullptr;  // device, batch entries
  int *ranges = nullptr; // device, ngroups + 1 entries
  // binding (owned plans): the pY contents the plan was built from, the
  // check kernel's flag and the generation of the last checked call
  void **ybind = nullptr; // device, batch pointers
  int *flag = nullptr;    // device, one int
  std::atomic<unsigned> gen{0};
};

struct km_group_bin
{
  int n;       // entry size of the bin
  int first;   // first range of the bin in binranges
  int start;   // first position of the bin in the plan's order
  int cnt;     // entries in the bin
  int chunk;   // entries per workgroup of the even (atomic) split
  int count;   // ranges of the bin
  int wide;    // FMA calls use the wider kron_lds (lds_kernel_wide)
};

struct km_group_plan_s
{
  int d = 0, batch = 0, elem_bytes = 0, nmax = 0;
  int device = 0;
  int ngroups = 0;
  int noutputs = 0;
  int *nvec = nullptr;   // device, per entry
  int *order = nullptr;  // device, batch entries
  // one-launch kernel (kron_runs): workgroup w takes work units
  // [ranges[w], ranges[w+1]); unit u covers order positions
  // [runstart[u], runstart[u+1]): a whole output run, or one part of a long
  // run of large entries split over several workgroups (unitmeta[3u..3u+2]
  // = split id, part, parts; split id -1 for a whole run).  Part sums of
  // split run s go to splitws + splitoff[s] + part*n^d; a second launch
  // (kron_runs_combine) adds them to the output in part order.
  int *ranges = nullptr;   // device, ngroups + 1 entries
  int *runstart = nullptr; // device, units + 1 entries
  int *unitmeta = nullptr; // device, 3 per unit
  void *splitws = nullptr; // device, partial sums of split runs
  long long *splitoff = nullptr; // device, per split run
  int *splitfirst = nullptr; // device, per split run: first order position
  int *splitparts = nullptr; // device, per split run: parts
  int nsplit = 0;
  // the units of small entries alone (one wavefront per entry in kron_runs),
  // cut into workgroup ranges: owned FMA calls of the per-size path run them
  // in one kron_runs launch instead of their size bins
  int *sranges = nullptr; // device, nsgroups + 1 entries (unit indices)
  int nsgroups = 0;
  int small_max = 0;      // largest n^d of a small entry (0: none)
  // per-size-bin launch: bin b uses binranges[first .. first+count]
  std::vector<km_group_bin> bins;
  int *binranges = nullptr; // device
  // binding of the owned launch modes (as km_plan_s)
  void **ybind = nullptr; // device, batch pointers
  int *flag = nullptr;    // device, one int
  std::atomic<unsigned> gen{0};
  // bins run concurrently on these streams (default priority), forked from
  // and joined back into the caller's stream with events
  std::vector<hipStream_t> streams;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  // held across a call's fork, launches and join: calls from several host
  // threads on one plan take turns with its streams and events
  std::mutex fork_mu;
};

namespace
{

std::mutex g_mu;

// detail of the last failure on this thread (km_last_error)
thread_local char g_last_error[256] = "";

void set_error(const char *where, hipError_t e)
{
  std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s (%d)", where,
                hipGetErrorString(e), (int)e);
}

// evaluate a HIP call; on failure record where and return KM_ERROR_HIP
#define KM_HIP(call)                                                           \
  do                                                                           \
  {                                                                            \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
    {                                                                          \
      set_error(#call, e_);                                                    \
      (void)hipGetLastError();                                                 \
      return KM_ERROR_HIP;                                                     \
    }                                                                          \
  } while (0)

// ---------------------------------------------------------------------------
// One HIP runtime per process.  A second copy of the runtime library (for
// instance PyTorch's own, loaded under another file name after this library
// bound the system one) keeps separate device state: pointers and streams of
// one are not valid in the other, and calls across them crash.  Entry points
// that touch HIP check for it first and return KM_ERROR_HIP instead.  The
// loaded objects are listed again only when the dynamic loader's counters of
// loads and unloads change.
// ---------------------------------------------------------------------------
struct RuntimeScan
{
  std::vector<std::pair<unsigned long, unsigned long>> files; // (dev, inode)
  std::string names;
};

int scan_runtime(struct dl_phdr_info *info, size_t, void *p)
{
  RuntimeScan *r = static_cast<RuntimeScan *>(p);
  const char *nm = info->dlpi_name;
  if (!nm || !*nm)
    return 0;
  const char *base = std::strrchr(nm, '/');
  base = base ? base + 1 : nm;
  if (std::strncmp(base, "libamdhip64.so", 14) != 0)
    return 0;
  struct stat st;
  if (stat(nm, &st) != 0)
    return 0;
  const std::pair<unsigned long, unsigned long> id((unsigned long)st.st_dev,
                                                   (unsigned long)st.st_ino);
  if (std::find(r->files.begin(), r->files.end(), id) == r->files.end())
  {
    r->files.push_back(id);
    r->names += r->names.empty() ? "" : ", ";
    r->names += nm;
  }
  return 0;
}

int loader_generation(struct dl_phdr_info *info, size_t size, void *p)
{
  unsigned long long *g = static_cast<unsigned long long *>(p);
  if (size >= offsetof(struct dl_phdr_info, dlpi_subs) + sizeof(info->dlpi_subs))
    *g = (unsigned long long)info->dlpi_adds * 1000003ull + info->dlpi_subs;
  return 1; // the counters are the same in every entry: stop at the first
}

} // namespace

namespace km
{
km_status runtime_check()
{
  static std::mutex mu;
  static unsigned long long seen = ~0ull;
  static bool two = false;
  static std::string names;
  unsigned long long gen = 0;
  dl_iterate_phdr(loader_generation, &gen);
  std::lock_guard<std::mutex> lk(mu);
  if (gen != seen || gen == 0)
  {
    RuntimeScan r;
    dl_iterate_phdr(scan_runtime, &r);
    two = r.files.size() > 1;
    names = r.names;
    seen = gen;
  }
  if (!two)
    return KM_SUCCESS;
  std::snprintf(g_last_error, sizeof(g_last_error),
                "two HIP runtimes are loaded in this process (%s): load the "
                "framework that owns the device buffers (e.g. PyTorch) before "
                "this library, so that both use one runtime",
                names.c_str());
  return KM_ERROR_HIP;
}
} // namespace km

namespace
{

int device_cus()
{
  static std::map<int, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return 0;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = cache.find(dev);
  if (it != cache.end())
    return it->second;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) !=
      hipSuccess)
    v = 0;
  cache[dev] = v;
  return v;
}

int occupancy_fn(const void *fn, int nt)
{
  static std::map<std::pair<int, const void *>, int> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = cache.find({dev, fn});
    if (it != cache.end())
      return it->second;
  }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, nt, 0) !=
          hipSuccess ||
      nb < 1)
    nb = 1;
  std::lock_guard<std::mutex> lk(g_mu);
  cache[{dev, fn}] = nb;
  return nb;
}

// workgroups per CU a launch of k is sized for: as many as fit, or fewer
// where the kernel records a measured limit (KernelInfo::max_per_cu)
int occupancy(const KernelInfo &k)
{
  const int nb = occupancy_fn(k.fn, k.nt);
  return k.max_per_cu > 0 ? std::min(nb, k.max_per_cu) : nb;
}

// kron_xcd runs on MI355X's layout (gfx950: eight XCDs, the CUs split evenly
// over them) and needs cooperative launches
bool xcd_device_ok()
{
  static std::map<int, bool> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return false;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = cache.find(dev);
    if (it != cache.end())
      return it->second;
  }
  hipDeviceProp_t prop;
  int coop = 0;
  const bool ok =
      hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 &&
      hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess &&
      coop != 0 && prop.multiProcessorCount >= 8 && prop.multiProcessorCount % 8 == 0;
  std::lock_guard<std::mutex> lk(g_mu);
  cache[dev] = ok;
  return ok;
}

long long ipow(int b, int e)
{
  long long p = 1;
  for (int i = 0; i < e; i++)
    p *= b;
  return p;
}

// Split positions [starts.front(), starts.back()) into at most G ranges that
// end only at run boundaries.  starts holds the first position of every run
// followed by the end position; pre[i] is the cost of the positions before
// starts[i] (pre[0] = 0, pre.back() = total).  Each cut is the boundary whose
// prefix cost is nearest k*total/G, so one resident wave of G workgroups
// covers all ranges and each range is within one run of the ideal cost.
// Appends the cuts and the end position to out (which ends with the start).
void balanced_cuts(const std::vector<int> &starts, const std::vector<double> &pre,
                   long long G, std::vector<int> &out)
{
  const double total = pre.back();
  for (long long k = 1; k < G; k++)
  {
    const double ideal = total * (double)k / (double)G;
    size_t idx = std::lower_bound(pre.begin(), pre.end(), ideal) - pre.begin();
    if (idx >= pre.size())
      idx = pre.size() - 1;
    if (idx > 0 && ideal - pre[idx - 1] < pre[idx] - ideal)
      idx--;
    const int cut = starts[idx];
    if (cut > out.back() && cut < starts.back())
      out.push_back(cut);
  }
  out.push_back(starts.back());
}

// kronm.hpp:37-89: forward only when strictly cheaper than backward
bool alg993_forward(int d, const int *m, const int *n)
{
  double xs = 1;
  for (int i = 0; i < d; i++)
    xs *= n[i];
  double fwd = 0, x = xs;
  for (int i = 0; i < d; i++)
  {
    double nn = x / n[i];
    fwd += 2.0 * m[i] * nn * n[i];
    x = m[i] * nn;
  }
  double bwd = 0;
  x = xs;
  for (int i = d - 1; i >= 0; i--)
  {
    double mm = x / n[i];
    bwd += 2.0 * mm * m[i] * n[i];
    x = mm * m[i];
  }
  return fwd < bwd;
}

struct Resolved
{
  // batch size and output ownership of the call, when known (selects the
  // tuned kernel variant; 0 = the built-in choice)
  long long batch = 0;
  int owned = 0;
  unsigned flags = 0;
  int precision = KM_PRECISION_FP64;
  bool exact = false;
  bool fwd = false;
  bool force_global = false;
  km_plan plan = nullptr;
};

km_status resolve(const km_options *opt, int d, const int *m, const int *n,
                  Resolved *r)
{
  km_options o;
  km_options_init(&o);
  if (opt)
    o = *opt;
  if (o.precision != KM_PRECISION_FP64 && o.precision != KM_PRECISION_FP64_MFMA &&
      o.precision != KM_PRECISION_F16X4)
    return KM_ERROR_NOT_SUPPORTED;
  r->precision = o.precision;
  r->flags = o.flags;
  r->exact = (o.flags & KM_FLAG_EXACT) != 0;
  r->force_global = (o.flags & KM_FLAG_FORCE_GLOBAL) != 0;
  switch (o.variant)
  {
  case KM_VARIANT_BACKWARD: r->fwd = false; break;
  case KM_VARIANT_FORWARD: r->fwd = true; break;
  case KM_VARIANT_ALG993: r->fwd = alg993_forward(d, m, n); break;
  default: return KM_ERROR_INVALID_VALUE;
  }
  r->plan = o.plan;
  return KM_SUCCESS;
}

// Square on-chip kernel for (d, n) honoring the rounding order, if any.
template<typename T>
bool pick_onchip(int d, int n, const Resolved &r, KernelInfo *k)
{
  if (r.force_global)
    return false;
  // FMA calls of a known batch size: the tuned variant, if any
  if (!r.exact && r.batch > 0)
  {
    const int v = km::tuned_variant(d, n, (int)sizeof(T), r.owned, r.batch);
    if (v != KM_KERNEL_DEFAULT && v != KM_KERNEL_TWOPASS && v != KM_KERNEL_XCD &&
        km::variant_kernel<T>(v, d, n, k))
      return true;
  }
  // without EXACT the contraction order only changes rounding, so the
  // forward request may use any kernel
  bool fwd = r.fwd && r.exact;
  // the register-blocked kernels contract in the forward order
  if (KM_USE_REG && (fwd || !r.exact) && km::reg_kernel<T>(d, n, r.exact, k))
    return true;
  if (km::lds_kernel<T>(d, n, r.exact, fwd, k))
    return true;
  if (!fwd && km::slab_kernel<T>(d, n, r.exact, k))
    return true;
  if (fwd && km::slab_fwd_kernel<T>(d, n, k))
    return true;
  return false;
}

// On-chip kernel of one size bin of a group plan (backward order when EXACT).
template<typename T>
bool bin_kernel(int d, int n, bool exact, KernelInfo *k)
{
  Resolved r;
  r.exact = exact;
  r.fwd = false;
  return pick_onchip<T>(d, n, r, k) &&
         (k->family == km::FAM_LDS || k->family == km::FAM_REG);
}

// A grouped size bin whose entries fit within this many resident waves of
// its kernel's workgroups is bound by one entry's chain of loads and passes:
// FMA calls served by kron_lds then use the wider kron_lds
// (lds_kernel_wide).  Measured on the grouped d=4, 20000-entry, n = 2..8
// call: 0.1505 -> 0.1413 ms with 4 (0.146 with 2, the same with 8).  Plain
// launches keep the narrow kernels: there the rule won on small batches
// (d5n5 x1000 0.026 -> 0.020 ms) but lost on others (d3n7 x5000 0.016 ->
// 0.021 ms, d4n5 x8000 0.028 -> 0.035 ms).
constexpr long long kWideWaves = 4;

// Whether a call's X vectors, each read once, are loaded with the
// non-temporal hint (EntryArgs::x_stream).  A call whose X fits well inside
// the 256 MB memory-side cache keeps the default policy: a repeated call
// finds X there.
constexpr long long kStreamBytes = 64ll << 20;
inline int x_streams(long long entries, long long xsize, int elem_bytes)
{
  return entries * xsize * elem_bytes >= kStreamBytes ? 1 : 0;
}

// Launch description common to all entry points.
template<typename T>
struct Call
{
  EntryArgs<T> a{};
  km::Shape sh{};
  bool square = true;
  long long xsize = 0; // elements of X per entry
  // workspace slots available to the staged kernel (0 = none)
  long long slots = 0;
  int elem_bytes = sizeof(T);
};

// Plan binding, for a call with an output-ordered plan and an output pointer
// array.  By default a check kernel compares the caller's pY with the
// pointers the plan was built from; where they differ, the call's kernels
// give up ownership and accumulate into the caller's pY with atomics
// (plan_bind).  With KM_FLAG_PLAN_OUTPUTS the call writes to the outputs the
// plan recorded and does not read pY.  Either way no output is updated by two
// workgroups without atomics.
template<typename T>
km_status bind_plan(Call<T> &c, const Resolved &r, hipStream_t stream)
{
  const km_plan p = r.plan;
  if (!p || !p->owned || !c.a.pY || !p->ybind)
    return KM_SUCCESS;
  if (r.flags & KM_FLAG_PLAN_OUTPUTS)
  {
    c.a.pY = reinterpret_cast<T *const *>(p->ybind);
    return KM_SUCCESS;
  }
  unsigned g = ++p->gen;
  if (g == 0) // the flag starts at 0
    g = ++p->gen;
  KM_HIP(km::plan_check_launch(reinterpret_cast<const void *const *>(c.a.pY),
                               p->ybind, c.a.batch, p->flag, g, stream));
  c.a.bind_flag = p->flag;
  c.a.bind_gen = g;
  return KM_SUCCESS;
}

// Two-pass family: pass 1 over the sub-slabs of a chunk of entries, pass 2
// over the chunk's entries.  Chunks are as large as the workspace allows:
// all entries when every entry brings its own workspace (batched,
// pbatched), else one slot per entry of the chunk (xbatched, vbatched).
template<typename T>
km_status launch_twopass(Call<T> &c, const Resolved &r, const KernelInfo &k1,
                         const KernelInfo &k2, hipStream_t stream)
{
  const int batch = c.a.batch;
  const int d = c.sh.d, n = c.sh.n[0];
  const int nn = n * n;
  const long long M = ipow(n, d - 2);
  const int cus = device_cus();
  if (cus <= 0)
    return KM_ERROR_HIP;
  const long long chunk_entries = c.a.slot_by_entry ? batch : c.slots;
  if (chunk_entries < 1)
    return KM_ERROR_INSUFFICIENT_WORKSPACE;
  if (r.plan)
  {
    const km_plan p = r.plan;
    if (p->batch != batch || p->elem_bytes != (int)sizeof(T) || p->d != d ||
        p->n != n)
      return KM_ERROR_PLAN_MISMATCH;
    if (chunk_entries < batch)
      return KM_ERROR_NOT_SUPPORTED; // a plan needs every entry's workspace
    c.a.order = p->order;
    km_status bs = bind_plan(c, r, stream);
    if (bs != KM_SUCCESS)
      return bs;
  }
  const int occ1 = occupancy(k1);
  const int occ2 = occupancy(k2);
  const int gy = (int)((M + k2.nt - 1) / k2.nt);
  for (long long e0 = 0; e0 < batch; e0 += chunk_entries)
  {
    const int cnt = (int)std::min<long long>(chunk_entries, batch - e0);
    int e0i = (int)e0;
    int e1i = (int)(e0 + cnt);
    // pass 1
    int nsub = cnt * nn;
    long long g1 = std::max(1LL, std::min<long long>((long long)cus * occ1, nsub));
    c.a.chunk = (int)((nsub + g1 - 1) / g1);
    c.a.ranges = nullptr;
    int grid1 = (nsub + c.a.chunk - 1) / c.a.chunk;
    void *args1[] = {&c.a, &e0i, &nsub};
    KM_HIP(hipLaunchKernel(k1.fn, dim3(grid1), dim3(k1.nt), args1, 0, stream));
    // pass 2
    int gx;
    if (r.plan)
    {
      c.a.ranges = r.plan->ranges;
      c.a.owned = r.plan->owned;
      gx = r.plan->ngroups;
    }
    else
    {
      long long want = std::max(1LL, (long long)cus * occ2 * 2 / gy);
      long long g = std::min<long long>(want, cnt);
      c.a.chunk = (int)((cnt + g - 1) / g);
      c.a.owned = 0;
      gx = (cnt + c.a.chunk - 1) / c.a.chunk;
    }
    void *args2[] = {&c.a, &e0i, &e1i};
    KM_HIP(hipLaunchKernel(k2.fn, dim3(gx, gy), dim3(k2.nt), args2, 0, stream));
  }
  return KM_SUCCESS;
}

// kron_xcd: the two-pass family in one cooperative launch, each XCD's W in
// its L2 (FMA calls whose entries each bring a workspace: batched,
// pbatched), where the variant table selects it (KM_KERNEL_XCD).  It moves
// half the HBM bytes of the two kernels but measured slower at every size
// tried (d6n8 x1024 2:1: 3.05 against 1.88 ms; d6n7 x16: 0.090 against
// 0.037 ms), so the two kernels stay the built-in choice.
// KM_ERROR_NOT_SUPPORTED leaves the call to the two kernels.  The spin limit
// bounds every barrier wait of the kernel (about a second).
constexpr unsigned kXcdSpinLimit = 1u << 20;

template<typename T>
km_status launch_xcd(Call<T> &c, const Resolved &r, hipStream_t stream)
{
  const int batch = c.a.batch;
  const int d = c.sh.d, n = c.sh.n[0];
  KernelInfo k;
  if (r.exact || !c.square || !c.a.slot_by_entry || c.slots < batch ||
      !km::xcd_kernel<T>(d, n, &k) || !xcd_device_ok())
    return KM_ERROR_NOT_SUPPORTED;
  if (km::tuned_variant(d, n, (int)sizeof(T), r.owned, batch) != KM_KERNEL_XCD)
    return KM_ERROR_NOT_SUPPORTED;
  // cooperative launches stay out of captured graphs
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return KM_ERROR_NOT_SUPPORTED;
  const int cus = device_cus();
  if (cus < 8)
    return KM_ERROR_NOT_SUPPORTED;
  if (r.plan)
  {
    const km_plan p = r.plan;
    if (p->batch != batch || p->elem_bytes != (int)sizeof(T) || p->d != d ||
        p->n != n)
      return KM_ERROR_PLAN_MISMATCH;
    c.a.order = p->order;
    c.a.owned = p->owned;
    km_status bs = bind_plan(c, r, stream);
    if (bs != KM_SUCCESS)
      return bs;
  }
  else
    c.a.owned = 0;
  c.a.ranges = nullptr;
  unsigned *ctl = nullptr;
  KM_HIP(km::xcd_prepare(stream, &ctl));
  unsigned limit = kXcdSpinLimit;
  void *args[] = {&c.a, &ctl, &limit};
  hipError_t e =
      hipLaunchCooperativeKernel(k.fn, dim3(cus * k.per_cu), dim3(k.nt), args, 0, stream);
  if (e == hipErrorCooperativeLaunchTooLarge)
  {
    (void)hipGetLastError();
    return KM_ERROR_NOT_SUPPORTED;
  }
  KM_HIP(e);
  return KM_SUCCESS;
}

// Forward two-pass family (exact, mode order 1..d): pass 1 over the chunk's
// entries and the positions of the trailing modes, pass 2 over the chunk's
// entries and the n*n sub-slabs.
template<typename T>
km_status launch_twopass_fwd(Call<T> &c, const Resolved &r, const KernelInfo &k1,
                             const KernelInfo &k2, hipStream_t stream)
{
  const int batch = c.a.batch;
  const int d = c.sh.d, n = c.sh.n[0];
  const int nn = n * n;
  const long long M = ipow(n, d - 2);
  const int cus = device_cus();
  if (cus <= 0)
    return KM_ERROR_HIP;
  const long long chunk_entries = c.a.slot_by_entry ? batch : c.slots;
  if (chunk_entries < 1)
    return KM_ERROR_INSUFFICIENT_WORKSPACE;
  if (r.plan)
  {
    const km_plan p = r.plan;
    if (p->batch != batch || p->elem_bytes != (int)sizeof(T) || p->d != d ||
        p->n != n)
      return KM_ERROR_PLAN_MISMATCH;
    if (chunk_entries < batch)
      return KM_ERROR_NOT_SUPPORTED; // a plan needs every entry's workspace
    c.a.order = p->order;
    km_status bs = bind_plan(c, r, stream);
    if (bs != KM_SUCCESS)
      return bs;
  }
  const int occ1 = occupancy(k1);
  const int occ2 = occupancy(k2);
  const int gy1 = (int)((M + k1.nt - 1) / k1.nt);
  for (long long e0 = 0; e0 < batch; e0 += chunk_entries)
  {
    const int cnt = (int)std::min<long long>(chunk_entries, batch - e0);
    int e0i = (int)e0;
    int e1i = (int)(e0 + cnt);
    // pass 1: entries x positions
    {
      long long want = std::max(1LL, (long long)cus * occ1 * 2 / gy1);
      long long g = std::min<long long>(want, cnt);
      c.a.chunk = (int)((cnt + g - 1) / g);
      c.a.ranges = nullptr;
      int gx = (cnt + c.a.chunk - 1) / c.a.chunk;
      void *args1[] = {&c.a, &e0i, &e1i};
      KM_HIP(hipLaunchKernel(k1.fn, dim3(gx, gy1), dim3(k1.nt), args1, 0, stream));
    }
    // pass 2: entries (plan ranges, or chunks) x sub-slabs
    int gx;
    if (r.plan)
    {
      c.a.ranges = r.plan->ranges;
      c.a.owned = r.plan->owned;
      gx = r.plan->ngroups;
    }
    else
    {
      long long want = std::max(1LL, (long long)cus * occ2 * 2 / nn);
      long long g = std::min<long long>(want, cnt);
      c.a.chunk = (int)((cnt + g - 1) / g);
      c.a.ranges = nullptr;
      c.a.owned = 0;
      gx = (cnt + c.a.chunk - 1) / c.a.chunk;
    }
    void *args2[] = {&c.a, &e0i, &e1i};
    KM_HIP(hipLaunchKernel(k2.fn, dim3(gx, nn), dim3(k2.nt), args2, 0, stream));
  }
  return KM_SUCCESS;
}

template<typename T>
km_status launch(Call<T> &c, const Resolved &r0, hipStream_t stream)
{
  const int batch = c.a.batch;
  if (batch == 0)
    return KM_SUCCESS;
  if (km_status g = km::runtime_check(); g != KM_SUCCESS)
    return g;
  Resolved r = r0;
  r.batch = batch;
  r.owned = (r.plan && r.plan->owned) ? 1 : 0;
  if (r.precision != KM_PRECISION_FP64)
  {
    // matrix-core pass 1, FP64 pass 2 (two-pass shapes, FMA calls only)
    KernelInfo k1, k2, kmx;
    if (!c.square || r.exact || r.force_global ||
        !km::twopass_kernels<T>(c.sh.d, c.sh.n[0], false, &k1, &k2) ||
        !km::twopass_mx_kernel<T>(c.sh.d, c.sh.n[0], r.precision, &kmx))
      return KM_ERROR_NOT_SUPPORTED;
    if (c.slots < 1)
      return KM_ERROR_INSUFFICIENT_WORKSPACE;
    return launch_twopass(c, r, kmx, k2, stream);
  }
  KernelInfo k;
  bool onchip = c.square && pick_onchip<T>(c.sh.d, c.sh.n[0], r, &k);
  if (!onchip && c.square && !r.force_global && !(r.fwd && r.exact))
  {
    KernelInfo k1, k2;
    if (km::twopass_kernels<T>(c.sh.d, c.sh.n[0], r.exact, &k1, &k2))
    {
      if (c.slots < 1)
        return KM_ERROR_INSUFFICIENT_WORKSPACE;
      const km_status xs = launch_xcd(c, r, stream);
      if (xs != KM_ERROR_NOT_SUPPORTED)
        return xs;
      return launch_twopass(c, r, k1, k2, stream);
    }
  }
  if (!onchip && c.square && !r.force_global && r.fwd && r.exact)
  {
    KernelInfo k1, k2;
    if (km::twopass_fwd_kernels<T>(c.sh.d, c.sh.n[0], &k1, &k2))
    {
      if (c.slots < 1)
        return KM_ERROR_INSUFFICIENT_WORKSPACE;
      return launch_twopass_fwd(c, r, k1, k2, stream);
    }
  }
  // the workspace-staged kernel caches fibers wider than KM_GLOBAL_MAX_N in
  // dynamic LDS, one slice of nmax values per thread: fewer threads per
  // workgroup for wider modes
  size_t global_lds = 0;
  if (!onchip)
  {
    if (c.slots < 1)
      return KM_ERROR_INSUFFICIENT_WORKSPACE;
    k = km::global_kernel<T>(r.exact);
    int nmax = 0;
    for (int i = 0; i < c.sh.d; i++)
      nmax = std::max(nmax, c.sh.n[i]);
    if (nmax > KM_GLOBAL_MAX_N)
    {
      const size_t per = sizeof(typename km::Acc<T>::type) * (size_t)nmax;
      const int nt = (int)std::min<size_t>(256, KM_GLOBAL_LDS_BYTES / per / 64 * 64);
      if (nt < 64)
        return KM_ERROR_NOT_SUPPORTED; // n_k above 300
      k.nt = nt;
      global_lds = per * nt;
    }
  }
  int cus = device_cus();
  if (cus <= 0)
  {
    std::snprintf(g_last_error, sizeof(g_last_error), "no device / CU count");
    return KM_ERROR_HIP;
  }
  // kron_half: two workgroups (a pair) per range of entries
  const bool halves = onchip && k.family == km::FAM_HALF;
  int grid = 0;
  if (r.plan)
  {
    const km_plan p = r.plan;
    if (p->batch != batch || p->elem_bytes != (int)sizeof(T) ||
        (c.square && (p->d != c.sh.d || p->n != c.sh.n[0])))
      return KM_ERROR_PLAN_MISMATCH;
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev != p->device)
      return KM_ERROR_PLAN_MISMATCH;
    grid = p->ngroups;
    if (!onchip && grid > c.slots)
      return KM_ERROR_INSUFFICIENT_WORKSPACE;
    c.a.ranges = p->ranges;
    c.a.order = p->order;
    c.a.owned = p->owned;
    km_status bs = bind_plan(c, r, stream);
    if (bs != KM_SUCCESS)
      return bs;
  }
  else
  {
    long long g = (long long)cus * occupancy(k);
    if (halves)
      g = std::max(1LL, g / 2);
    if (!onchip)
      g = std::min(g, c.slots);
    g = std::max(1LL, std::min<long long>(g, batch));
    int chunk = (int)((batch + g - 1) / g);
    grid = (batch + chunk - 1) / chunk;
    c.a.chunk = chunk;
    c.a.ranges = nullptr;
    c.a.order = nullptr;
    c.a.owned = 0;
  }
  hipError_t e;
  if (halves)
  {
    // pair q is served by workgroups 16*(q/8) + q%8 and that + 8
    int npairs = grid;
    const int g2 = (2 * npairs + 15) / 16 * 16;
    void *args[] = {&c.a, &npairs};
    e = hipLaunchKernel(k.fn, dim3(g2), dim3(k.nt), args, 0, stream);
  }
  else if (onchip)
  {
    void *args[] = {&c.a};
    e = hipLaunchKernel(k.fn, dim3(grid), dim3(k.nt), args, 0, stream);
  }
  else
  {
    int fwd = r.fwd ? 1 : 0;
    void *args[] = {&c.a, &c.sh, &fwd};
    e = hipLaunchKernel(k.fn, dim3(grid), dim3(k.nt), args, global_lds, stream);
  }
  if (e != hipSuccess)
  {
    set_error("hipLaunchKernel", e);
    (void)hipGetLastError();
    return KM_ERROR_HIP;
  }
  return KM_SUCCESS;
}

km_status check_dn(int d, int n, int batch)
{
  if (d < 1 || d > 6 || n < 1 || batch < 0)
    return KM_ERROR_INVALID_VALUE;
  if (ipow(n, d) > (1LL << 31) - 1)
    return KM_ERROR_INVALID_VALUE;
  return KM_SUCCESS;
}

void square_shape(km::Shape &sh, int d, int n)
{
  sh.d = d;
  for (int k = 0; k < 6; k++)
  {
    sh.m[k] = (k < d) ? n : 1;
    sh.n[k] = (k < d) ? n : 1;
  }
}

template<typename T>
km_status batched(int d, int n, const T *A, const T *X, T *Y, T *W, int batch,
                  void *stream, const km_options *opt)
{
  km_status s = check_dn(d, n, batch);
  if (s != KM_SUCCESS)
    return s;
  if (batch > 0 && (!A || !X || !Y))
    return KM_ERROR_INVALID_VALUE;
  Call<T> c;
  square_shape(c.sh, d, n);
  Resolved r;
  if ((s = resolve(opt, d, c.sh.m, c.sh.n, &r)) != KM_SUCCESS)
    return s;
  const long long xs = ipow(n, d);
  c.xsize = xs;
  c.a.A = A;
  c.a.a_stride = (long long)n * n * d;
  c.a.a_kstride = (long long)n * n;
  c.a.lda = n;
  c.a.X = X;
  c.a.x_stride = xs;
  c.a.Y = Y;
  c.a.y_stride = xs;
  c.a.W = W;
  c.a.w_stride = xs;
  c.a.w_half = 0;
  c.a.ws_mul = 1;
  c.a.slot_by_entry = 1;
  c.a.x_stream = x_streams(batch, xs, sizeof(T));
  c.a.batch = batch;
  c.slots = W ? batch : 0;
  return launch(c, r, (hipStream_t)stream);
}

template<typename T>
km_status pbatched(int d, int n, const T *A, T *const *pX, T *const *pY,
                   T *const *pW, int batch, void *stream,
                   const km_options *opt)
{
  km_status s = check_dn(d, n, batch);
  if (s != KM_SUCCESS)
    return s;
  if (batch > 0 && (!A || !pX || !pY))
    return KM_ERROR_INVALID_VALUE;
  Call<T> c;
  square_shape(c.sh, d, n);
  Resolved r;
  if ((s = resolve(opt, d, c.sh.m, c.sh.n, &r)) != KM_SUCCESS)
    return s;
  c.xsize = ipow(n, d);
  c.a.A = A;
  c.a.a_stride = (long long)n * n * d;
  c.a.a_kstride = (long long)n * n;
  c.a.lda = n;
  c.a.pX = pX;
  c.a.pY = pY;
  c.a.pW = pW;
  c.a.ws_mul = 1;
  c.a.slot_by_entry = 1;
  c.a.x_stream = x_streams(batch, c.xsize, sizeof(T));
  c.a.batch = batch;
  c.slots = pW ? batch : 0;
  return launch(c, r, (hipStream_t)stream);
}

template<typename T>
km_status xbatched(int d, int n, const T *const *Aarray, int lda, T *const *pX,
                   T *const *pY, T *const *pW, int batch, int sub,
                   void *stream, const km_options *opt)
{
  km_status s = check_dn(d, n, batch);
  if (s != KM_SUCCESS)
    return s;
  if (lda < n || sub < 0)
    return KM_ERROR_INVALID_VALUE;
  if (batch > 0 && (!Aarray || !pX || !pY))
    return KM_ERROR_INVALID_VALUE;
  Call<T> c;
  square_shape(c.sh, d, n);
  Resolved r;
  if ((s = resolve(opt, d, c.sh.m, c.sh.n, &r)) != KM_SUCCESS)
    return s;
  if (sub == 0)
    sub = std::max(1, batch / 2); // kronmult6_xbatched.cpp:16-20
  c.xsize = ipow(n, d);
  c.a.Aptr = Aarray;
  c.a.lda = lda;
  c.a.pX = pX;
  c.a.pY = pY;
  c.a.pW = pW;
  c.a.ws_mul = 2; // slot s uses the pair pW[2s], pW[2s+1]
  c.a.x_stream = 0; // pooled inputs: entries share X vectors
  c.a.batch = batch;
  c.slots = pW ? sub : 0;
  return launch(c, r, (hipStream_t)stream);
}

template<typename T>
km_status vbatched(int d, const int *m, const int *n, const T *const *Aarray,
                   T *const *pX, T *const *pY, T *W, size_t Wcap, int batch,
                   void *stream, const km_options *opt)
{
  if (d < 1 || d > 6 || batch < 0 || !m || !n)
    return KM_ERROR_INVALID_VALUE;
  for (int k = 0; k < d; k++)
    if (m[k] < 1 || n[k] < 1)
      return KM_ERROR_INVALID_VALUE;
  if (batch > 0 && (!Aarray || !pX || !pY))
    return KM_ERROR_INVALID_VALUE;
  Call<T> c;
  c.sh.d = d;
  bool square = true;
  for (int k = 0; k < 6; k++)
  {
    c.sh.m[k] = (k < d) ? m[k] : 1;
    c.sh.n[k] = (k < d) ? n[k] : 1;
    if (k < d)
      square = square && (m[k] == n[0] && n[k] == n[0]);
  }
  c.square = square;
  Resolved r;
  km_status s;
  if ((s = resolve(opt, d, c.sh.m, c.sh.n, &r)) != KM_SUCCESS)
    return s;
  // Largest intermediate of the stage order in use (for the forward order
  // this is the reference's sizeW, kronmult_vbatched.hpp:114-130; the
  // backward order contracts the last modes first).
  long long sx = 1, sw = 0;
  for (int k = 0; k < d; k++)
    sx *= n[k];
  for (int L = 1; L < d; L++)
  {
    long long v = 1;
    for (int k = 0; k < d; k++)
    {
      bool done = r.fwd ? (k < L) : (k >= d - L);
      v *= done ? m[k] : n[k];
    }
    sw = std::max(sw, v);
  }
  long long sy = 1;
  for (int k = 0; k < d; k++)
    sy *= m[k];
  if (sx > (1LL << 31) - 1 || sy > (1LL << 31) - 1 || sw > (1LL << 31) - 1)
    return KM_ERROR_INVALID_VALUE;
  const long long sxw = std::max(std::max(sx, sw), sy);
  c.xsize = sx;
  c.a.Aptr = Aarray;
  c.a.lda = square ? m[0] : 0; // 0: A_k has leading dimension m_k
  c.a.pX = pX;
  c.a.pY = pY;
  c.a.W = W;
  c.a.w_stride = 2 * sxw;
  c.a.w_half = sxw;
  c.a.x_stream = x_streams(batch, sx, sizeof(T));
  c.a.batch = batch;
  c.slots = W ? (long long)(Wcap / (2 * sxw * sizeof(T))) : 0;
  return launch(c, r, (hipStream_t)stream);
}

template<typename T>
km_status grouped(km_group_plan plan, const T *const *Aarray, T *const *pX,
                  T *const *pY, void *stream, const km_options *opt)
{
  if (!plan)
    return KM_ERROR_INVALID_VALUE;
  if (plan->elem_bytes != (int)sizeof(T))
    return KM_ERROR_PLAN_MISMATCH;
  if (plan->batch == 0)
    return KM_SUCCESS;
  if (!Aarray || !pX || !pY)
    return KM_ERROR_INVALID_VALUE;
  if (km_status g = km::runtime_check(); g != KM_SUCCESS)
    return g;
  int dev = 0;
  (void)hipGetDevice(&dev);
  if (dev != plan->device)
    return KM_ERROR_PLAN_MISMATCH;
  km_options o;
  km_options_init(&o);
  if (opt)
    o = *opt;
  if (o.precision != KM_PRECISION_FP64 || o.plan)
    return KM_ERROR_NOT_SUPPORTED;
  const bool exact = (o.flags & KM_FLAG_EXACT) != 0;
  // owned launch modes are bound to the plan's outputs (see bind_plan): a
  // check kernel on the caller's stream ahead of every launch (and of the
  // fork), or the plan's recorded outputs with KM_FLAG_PLAN_OUTPUTS
  const bool owned_mode =
      (o.flags & (KM_FLAG_PERSISTENT | KM_FLAG_DETERMINISTIC)) != 0;
  const int *bind_flag = nullptr;
  unsigned bind_gen = 0;
  if (owned_mode && plan->ybind)
  {
    if (o.flags & KM_FLAG_PLAN_OUTPUTS)
      pY = reinterpret_cast<T *const *>(plan->ybind);
    else
    {
      bind_gen = ++plan->gen;
      if (bind_gen == 0) // the flag starts at 0
        bind_gen = ++plan->gen;
      KM_HIP(km::plan_check_launch(reinterpret_cast<const void *const *>(pY),
                                   plan->ybind, plan->batch, plan->flag,
                                   bind_gen, (hipStream_t)stream));
      bind_flag = plan->flag;
    }
  }
  if (!(o.flags & KM_FLAG_PERSISTENT))
  {
    // one LDS-resident launch per size bin, each over its plan ranges
    EntryArgs<T> a{};
    a.Aptr = Aarray;
    a.lda = 0;
    a.pX = pX;
    a.pY = pY;
    a.bind_flag = bind_flag;
    a.bind_gen = bind_gen;
    long long xtotal = 0;
    for (const km_group_bin &bn : plan->bins)
      xtotal += (long long)bn.cnt * ipow(bn.n, plan->d);
    a.x_stream = x_streams(1, xtotal, sizeof(T));
    const bool owned = (o.flags & KM_FLAG_DETERMINISTIC) != 0;
    // owned FMA calls: the size bins of small entries (one wavefront per
    // entry in kron_runs) as one kron_runs launch over the plan's ranges of
    // their output runs, beside the larger bins.  Measured on the d = 4,
    // n = 2..8 mix: 0.226 -> 0.186 ms; with n = 2..5 alone 0.179 -> 0.044 ms.
    // The default (atomic) calls keep their bins: there one launch per run
    // caps the parallelism at the number of outputs (n = 2 alone with 285
    // outputs: 9 -> 33 us), and the 2..8 mix was unchanged (0.138 ms).
    KernelInfo kr;
    int rn = 0, rsm = 0;
    const bool runs_small = owned && !exact && plan->nsgroups > 0 &&
                            km::runs_kernel<T>(plan->d, false, &kr, &rn, &rsm) &&
                            plan->nmax <= rn && rsm == plan->small_max;
    auto small_bin = [&](const km_group_bin &bn) {
      return runs_small && ipow(bn.n, plan->d) <= plan->small_max;
    };
    // resolve every bin's kernel before the fork, so an unsupported size
    // returns with nothing launched and no stream left forked
    std::vector<KernelInfo> kbs(plan->bins.size());
    for (size_t i = 0; i < plan->bins.size(); i++)
    {
      const km_group_bin &bn = plan->bins[i];
      if (small_bin(bn))
        continue;
      // owned FMA bins take every kron_reg built (a bin of whole output runs
      // has few, long workgroups: d4n5 and d4n7 bins of 2857 entries measured
      // 0.018 and 0.057 ms against 0.032 and 0.073 with kron_lds)
      const bool ok = (owned && !exact && km::reg_kernel<T>(plan->d, bn.n, false, &kbs[i], true)) ||
                      ((bn.wide && !exact) ? km::lds_kernel_wide<T>(plan->d, bn.n, &kbs[i])
                                           : bin_kernel<T>(plan->d, bn.n, exact, &kbs[i]));
      if (!ok)
        return KM_ERROR_NOT_SUPPORTED;
    }
    // owned outputs: a bin's grid is its number of output runs, often far
    // below a full chip, so the bins run concurrently on the plan's streams.
    // The even split of the atomic path already sizes each bin's grid to
    // fill the chip, and there the fork and join only add latency.
    // Only as many streams as launches are forked (none for one launch: a
    // lone kron_runs launch on four forked streams took 0.104 ms where it
    // takes 0.044 on the caller's stream).
    int nlaunch = runs_small ? 1 : 0;
    for (const km_group_bin &bn : plan->bins)
      nlaunch += small_bin(bn) ? 0 : 1;
    const int ns = owned ? std::min<int>((int)plan->streams.size(), nlaunch) : 0;
    std::unique_lock<std::mutex> fork_lock(plan->fork_mu, std::defer_lock);
    if (ns > 1)
    {
      fork_lock.lock();
      KM_HIP(hipEventRecord(plan->fork, (hipStream_t)stream));
      for (int i = 0; i < ns; i++)
        KM_HIP(hipStreamWaitEvent(plan->streams[i], plan->fork, 0));
    }
    hipError_t launched = hipSuccess;
    int nl = 0; // launches so far (dealt out to the streams in turn)
    for (size_t bi = 0; bi < plan->bins.size(); bi++)
    {
      const km_group_bin &bn = plan->bins[bi];
      if (small_bin(bn))
        continue;
      const KernelInfo &kb = kbs[bi];
      hipStream_t bs = ns > 1 ? plan->streams[nl++ % ns] : (hipStream_t)stream;
      int grid;
      if (owned)
      {
        // ranges of whole output runs, positions into the full order
        a.batch = plan->batch;
        a.order = plan->order;
        a.ranges = plan->binranges + bn.first;
        a.chunk = 0;
        a.owned = 1;
        grid = bn.count;
      }
      else
      {
        // even split of the bin's positions; equal outputs stay adjacent
        // (merged on chip), the run shared by two workgroups uses atomics
        a.batch = bn.cnt;
        a.order = plan->order + bn.start;
        a.ranges = nullptr;
        a.chunk = bn.chunk;
        a.owned = 0;
        grid = (bn.cnt + bn.chunk - 1) / bn.chunk;
      }
      void *args[] = {&a};
      launched = hipLaunchKernel(kb.fn, dim3(grid), dim3(kb.nt), args, 0, bs);
      if (launched != hipSuccess)
        break;
    }
    if (runs_small && launched == hipSuccess)
    {
      EntryArgs<T> ar = a;
      ar.batch = plan->batch;
      ar.order = plan->order;
      ar.ranges = plan->sranges;
      ar.chunk = 0;
      ar.owned = 1;
      km::RunsArgs<T> ra;
      ra.nvec = plan->nvec;
      ra.ustart = plan->runstart;
      ra.umeta = plan->unitmeta;
      ra.ws = static_cast<T *>(plan->splitws);
      ra.wsoff = plan->splitoff;
      ra.sfirst = plan->splitfirst;
      ra.sparts = plan->splitparts;
      void *args[] = {&ar, &ra};
      hipStream_t rs = ns > 1 ? plan->streams[nl++ % ns] : (hipStream_t)stream;
      launched = hipLaunchKernel(kr.fn, dim3(plan->nsgroups), dim3(kr.nt), args, 0, rs);
    }
    // the join also follows a failed launch, so that the caller's stream
    // (or a graph being captured from it) is never left forked
    if (ns > 1)
    {
      for (int i = 0; i < ns; i++)
      {
        KM_HIP(hipEventRecord(plan->joins[i], plan->streams[i]));
        KM_HIP(hipStreamWaitEvent((hipStream_t)stream, plan->joins[i], 0));
      }
    }
    KM_HIP(launched);
    return KM_SUCCESS;
  }
  KernelInfo k;
  int nmax = 0, small_max = 0;
  if (!km::runs_kernel<T>(plan->d, exact, &k, &nmax, &small_max) || plan->nmax > nmax)
    return KM_ERROR_NOT_SUPPORTED;
  EntryArgs<T> a{};
  a.Aptr = Aarray;
  a.lda = 0; // A_k of entry b has leading dimension n_b
  a.pX = pX;
  a.pY = pY;
  a.batch = plan->batch;
  a.ranges = plan->ranges;
  a.order = plan->order;
  a.owned = 1;
  a.bind_flag = bind_flag;
  a.bind_gen = bind_gen;
  km::RunsArgs<T> ra;
  ra.nvec = plan->nvec;
  ra.ustart = plan->runstart;
  ra.umeta = plan->unitmeta;
  ra.ws = static_cast<T *>(plan->splitws);
  ra.wsoff = plan->splitoff;
  ra.sfirst = plan->splitfirst;
  ra.sparts = plan->splitparts;
  void *args[] = {&a, &ra};
  KM_HIP(hipLaunchKernel(k.fn, dim3(plan->ngroups), dim3(k.nt), args, 0,
                         (hipStream_t)stream));
  if (plan->nsplit > 0)
  {
    // the parts of split runs, added to their outputs in part order
    int cnt_nt = 0;
    const void *cf = km::runs_combine_kernel<T>(plan->d, &cnt_nt);
    if (!cf)
      return KM_ERROR_NOT_SUPPORTED;
    KM_HIP(hipLaunchKernel(cf, dim3(plan->nsplit), dim3(cnt_nt), args, 0,
                           (hipStream_t)stream));
  }
  return KM_SUCCESS;
}

template<typename T>
km_status kgemm(int op, int mm, int nn, int kk, T alpha, const T *const *A,
                const int *ldA, const T *const *B, const int *ldB, T beta,
                T *const *C, const int *ldC, int batch, void *stream,
                const km_options *opt)
{
  if (mm < 0 || nn < 0 || kk < 0 || batch < 0)
    return KM_ERROR_INVALID_VALUE;
  if (mm == 0 || nn == 0 || batch == 0)
    return KM_SUCCESS;
  if (!A || !ldA || !B || !ldB || !C || !ldC)
    return KM_ERROR_INVALID_VALUE;
  km_options o;
  km_options_init(&o);
  if (opt)
    o = *opt;
  if (o.precision != KM_PRECISION_FP64 || o.plan)
    return KM_ERROR_NOT_SUPPORTED;
  if (km_status rc = km::runtime_check(); rc != KM_SUCCESS)
    return rc;
  const int cus = device_cus();
  if (cus <= 0)
  {
    std::snprintf(g_last_error, sizeof(g_last_error), "no device / CU count");
    return KM_ERROR_HIP;
  }
  km::KgemmArgs<T> g{};
  g.A = A;
  g.B = B;
  g.C = C;
  g.lda = ldA;
  g.ldb = ldB;
  g.ldc = ldC;
  g.mm = mm;
  g.nn = nn;
  g.kk = kk;
  g.batch = batch;
  g.alpha = alpha;
  g.beta = beta;
  KM_HIP(km::kgemm_launch<T>(op, g, (o.flags & KM_FLAG_EXACT) != 0, cus,
                             (hipStream_t)stream));
  return KM_SUCCESS;
}

template<typename T>
km_status xbatched_tables(int d, int batch, const T *const *Apool,
                          const long long *a_stride, const int *a_count,
                          const int *aidx, T *X, long long x_stride, int x_count,
                          const int *xidx, T *Y, long long y_stride, int y_count,
                          const int *yidx, const T **Aarray, T **pX, T **pY,
                          int *bad, void *stream)
{
  if (d < 1 || d > 6 || batch < 0)
    return KM_ERROR_INVALID_VALUE;
  if (batch == 0)
    return KM_SUCCESS;
  if (!Apool || !a_stride || !a_count || !aidx || !X || !xidx || !Y || !yidx ||
      !Aarray || !pX || !pY || x_stride < 0 || y_stride < 0 || x_count < 0 ||
      y_count < 0)
    return KM_ERROR_INVALID_VALUE;
  if (km_status rc = km::runtime_check(); rc != KM_SUCCESS)
    return rc;
  km::TableArgs<T> t{};
  t.d = d;
  t.batch = batch;
  for (int k = 0; k < d; k++)
  {
    if (!Apool[k] || a_stride[k] < 0 || a_count[k] < 0)
      return KM_ERROR_INVALID_VALUE;
    t.pool[k] = Apool[k];
    t.a_stride[k] = a_stride[k];
    t.a_count[k] = a_count[k];
  }
  t.aidx = aidx;
  t.xidx = xidx;
  t.yidx = yidx;
  t.X = X;
  t.x_stride = x_stride;
  t.x_count = x_count;
  t.Y = Y;
  t.y_stride = y_stride;
  t.y_count = y_count;
  t.Aarray = Aarray;
  t.pX = pX;
  t.pY = pY;
  t.bad = bad;
  const int cus = device_cus();
  KM_HIP(km::tables_launch<T>(t, cus, (hipStream_t)stream));
  return KM_SUCCESS;
}

// km_kronmult_pbatched_noalloc_f64: pbatched with an output-ordered plan that
// lives on the stack and points into the static per-device buffer, where the
// order is built on the device; no allocation, no host synchronization.
template<typename T>
km_status pbatched_noalloc(int d, int n, const T *A, T *const *pX,
                           T *const *pY, T *const *pW, int batch, void *stream,
                           const km_options *opt);

// Workgroup count of an output-ordered plan for a square (d, n) call: one
// resident wave of the default kernel's workgroups (of pass 2 for the
// two-pass family), at most one per entry.
long long plan_groups(int d, int n, int elem_bytes, int batch)
{
  Resolved r;
  r.batch = batch;
  r.owned = 1;
  KernelInfo k;
  bool onchip = (elem_bytes == 8) ? pick_onchip<double>(d, n, r, &k)
                                  : pick_onchip<float>(d, n, r, &k);
  long long G;
  KernelInfo k1, k2;
  bool twopass = !onchip && ((elem_bytes == 8)
                                 ? km::twopass_kernels<double>(d, n, false, &k1, &k2)
                                 : km::twopass_kernels<float>(d, n, false, &k1, &k2));
  if (!onchip && !twopass)
    k = (elem_bytes == 8) ? km::global_kernel<double>(false)
                          : km::global_kernel<float>(false);
  if (twopass)
  {
    // pass 2 runs one workgroup per (entry group, block of positions)
    const long long gy = (ipow(n, d - 2) + k2.nt - 1) / k2.nt;
    G = std::max(1LL, (long long)device_cus() * occupancy(k2) * 2 / gy);
  }
  else
    G = (long long)device_cus() * occupancy(k);
  if (onchip && k.family == km::FAM_HALF)
    G = std::max(1LL, G / 2); // ranges are served by pairs of workgroups
  G = std::min<long long>(G, km::kOrderMaxGroups);
  return std::max(1LL, std::min<long long>(G, batch));
}

template<typename T>
km_status pbatched_noalloc(int d, int n, const T *A, T *const *pX,
                           T *const *pY, T *const *pW, int batch, void *stream,
                           const km_options *opt)
{
  km_status s = check_dn(d, n, batch);
  if (s != KM_SUCCESS)
    return s;
  if (opt && opt->plan)
    return KM_ERROR_INVALID_VALUE;
  if (batch > KM_NOALLOC_MAX_BATCH)
  {
    std::snprintf(g_last_error, sizeof(g_last_error),
                  "km_kronmult_pbatched_noalloc: batchCount %d exceeds "
                  "KM_NOALLOC_MAX_BATCH (%d)",
                  batch, KM_NOALLOC_MAX_BATCH);
    return KM_ERROR_INVALID_VALUE;
  }
  if (batch == 0)
    return KM_SUCCESS;
  if (!A || !pX || !pY)
    return KM_ERROR_INVALID_VALUE;
  const int G = (int)plan_groups(d, n, sizeof(T), batch);
  if (km::order_bytes(batch, G) > KM_NOALLOC_SCRATCH_BYTES)
  {
    std::snprintf(g_last_error, sizeof(g_last_error),
                  "km_kronmult_pbatched_noalloc: order of %d entries in %d "
                  "ranges exceeds the static buffer",
                  batch, G);
    return KM_ERROR_INVALID_VALUE;
  }
  if (km_status g = km::runtime_check(); g != KM_SUCCESS)
    return g;
  void *buf = nullptr;
  KM_HIP(km::noalloc_scratch(&buf));
  km::OrderOut o{};
  KM_HIP(km::order_build(reinterpret_cast<const void *const *>(pY), batch, G,
                         buf, (hipStream_t)stream, &o));
  // a plan for this call only: G workgroups, the empty ranges at the end
  // exit at once; no binding copy (the order was built from this pY)
  km_plan_s p;
  p.d = d;
  p.n = n;
  p.elem_bytes = (int)sizeof(T);
  p.batch = batch;
  KM_HIP(hipGetDevice(&p.device));
  p.ngroups = G;
  p.noutputs = 0;
  p.owned = 1;
  p.order = o.order;
  p.ranges = o.ranges;
  km_options o2;
  km_options_init(&o2);
  if (opt)
    o2 = *opt;
  o2.plan = &p;
  return pbatched<T>(d, n, A, pX, pY, pW, batch, stream, &o2);
}

} // namespace

// ===========================================================================
// extern "C"
// ===========================================================================
extern "C"
{

  void km_options_init(km_options *opt)
  {
    if (!opt)
      return;
    std::memset(opt, 0, sizeof(*opt));
    opt->flags = 0;
    opt->variant = KM_VARIANT_BACKWARD;
    opt->precision = KM_PRECISION_FP64;
    opt->plan = nullptr;
  }

  const char *km_last_error(void) { return g_last_error; }

  const char *km_status_string(km_status s)
  {
    switch (s)
    {
    case KM_SUCCESS: return "success";
    case KM_ERROR_INVALID_VALUE: return "invalid value";
    case KM_ERROR_INSUFFICIENT_WORKSPACE: return "insufficient workspace";
    case KM_ERROR_NOT_SUPPORTED: return "not supported";
    case KM_ERROR_HIP: return "HIP runtime error";
    case KM_ERROR_PLAN_MISMATCH: return "plan does not match the call";
    }
    return "unknown status";
  }

  km_status km_kronmult_batched_f64(int d, int n, const double *A,
                                    const double *X, double *Y, double *W,
                                    int batchCount, void *stream,
                                    const km_options *opt)
  {
    return batched<double>(d, n, A, X, Y, W, batchCount, stream, opt);
  }
  km_status km_kronmult_batched_f32(int d, int n, const float *A,
                                    const float *X, float *Y, float *W,
                                    int batchCount, void *stream,
                                    const km_options *opt)
  {
    return batched<float>(d, n, A, X, Y, W, batchCount, stream, opt);
  }

  km_status km_kronmult_pbatched_f64(int d, int n, const double *A,
                                     double *const *pX, double *const *pY,
                                     double *const *pW, int batchCount,
                                     void *stream, const km_options *opt)
  {
    return pbatched<double>(d, n, A, pX, pY, pW, batchCount, stream, opt);
  }
  km_status km_kronmult_pbatched_f32(int d, int n, const float *A,
                                     float *const *pX, float *const *pY,
                                     float *const *pW, int batchCount,
                                     void *stream, const km_options *opt)
  {
    return pbatched<float>(d, n, A, pX, pY, pW, batchCount, stream, opt);
  }

  km_status km_kronmult_xbatched_f64(int d, int n,
                                     const double *const *Aarray, int lda,
                                     double *const *pX, double *const *pY,
                                     double *const *pW, int batchCount,
                                     int subbatchCount, void *stream,
                                     const km_options *opt)
  {
    return xbatched<double>(d, n, Aarray, lda, pX, pY, pW, batchCount,
                            subbatchCount, stream, opt);
  }
  km_status km_kronmult_xbatched_f32(int d, int n, const float *const *Aarray,
                                     int lda, float *const *pX,
                                     float *const *pY, float *const *pW,
                                     int batchCount, int subbatchCount,
                                     void *stream, const km_options *opt)
  {
    return xbatched<float>(d, n, Aarray, lda, pX, pY, pW, batchCount,
                           subbatchCount, stream, opt);
  }

  km_status km_kronmult_vbatched_f64(int d, const int *m, const int *n,
                                     const double *const *Aarray,
                                     double *const *pX, double *const *pY,
                                     double *W, size_t Wcapacity_bytes,
                                     int batchCount, void *stream,
                                     const km_options *opt)
  {
    return vbatched<double>(d, m, n, Aarray, pX, pY, W, Wcapacity_bytes,
                            batchCount, stream, opt);
  }
  km_status km_kronmult_vbatched_f32(int d, const int *m, const int *n,
                                     const float *const *Aarray,
                                     float *const *pX, float *const *pY,
                                     float *W, size_t Wcapacity_bytes,
                                     int batchCount, void *stream,
                                     const km_options *opt)
  {
    return vbatched<float>(d, m, n, Aarray, pX, pY, W, Wcapacity_bytes,
                           batchCount, stream, opt);
  }

  km_status km_plan_create(km_plan *plan, int d, int n, int elem_bytes,
                           const void *const *pY, int batchCount,
                           void *stream)
  {
    return km_plan_create_ordered(plan, d, n, elem_bytes, nullptr, pY,
                                  batchCount, KM_ORDER_BY_OUTPUT, stream);
  }

  km_status km_plan_create_ordered(km_plan *plan, int d, int n, int elem_bytes,
                                   const void *const *pX,
                                   const void *const *pY, int batchCount,
                                   int order_by, void *stream)
  {
    if (!plan || (elem_bytes != 4 && elem_bytes != 8) || batchCount < 1 ||
        !pY ||
        (order_by != KM_ORDER_BY_OUTPUT && order_by != KM_ORDER_BY_INPUT) ||
        (order_by == KM_ORDER_BY_INPUT && !pX))
      return KM_ERROR_INVALID_VALUE;
    *plan = nullptr;
    km_status s = check_dn(d, n, batchCount);
    if (s != KM_SUCCESS)
      return s;
    if ((s = km::runtime_check()) != KM_SUCCESS)
      return s;
    hipStream_t st = (hipStream_t)stream;
    const long long G = plan_groups(d, n, elem_bytes, batchCount);
    km_plan p = new km_plan_s;
    p->owned = (order_by == KM_ORDER_BY_OUTPUT) ? 1 : 0;
    p->d = d;
    p->n = n;
    p->elem_bytes = elem_bytes;
    p->batch = batchCount;
    (void)hipGetDevice(&p->device);
    km_status st2 = [&]() -> km_status {
      if (order_by == KM_ORDER_BY_OUTPUT && batchCount <= km::kOrderMaxBatch)
      {
        // built on the device: stable radix sort by output, run starts and
        // balanced cuts (km_order.hip); only the output count comes back
        void *tmp = nullptr;
        KM_HIP(hipMalloc(&tmp, km::order_bytes(batchCount, (int)G)));
        km::OrderOut o{};
        hipError_t e = km::order_build(pY, batchCount, (int)G, tmp, st, &o);
        int nout[2] = {0, 0};
        if (e == hipSuccess)
          e = hipMalloc(&p->order, sizeof(int) * batchCount);
        if (e == hipSuccess)
          e = hipMalloc(&p->ranges, sizeof(int) * (G + 1));
        if (e == hipSuccess)
          e = hipMemcpyAsync(p->order, o.order, sizeof(int) * batchCount,
                             hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
          e = hipMemcpyAsync(p->ranges, o.ranges, sizeof(int) * (G + 1),
                             hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
          e = hipMemcpyAsync(nout, o.nout, sizeof(nout), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
          e = hipStreamSynchronize(st);
        (void)hipFree(tmp);
        KM_HIP(e);
        // launches cover the non-empty ranges only
        p->ngroups = nout[1];
        p->noutputs = nout[0];
      }
      else
      {
        // host-built order: input-ordered plans, and batches beyond the
        // device builder's limit
        std::vector<uintptr_t> ys(batchCount), xs;
        KM_HIP(hipMemcpyAsync(ys.data(), pY, sizeof(void *) * batchCount,
                              hipMemcpyDeviceToHost, st));
        if (order_by == KM_ORDER_BY_INPUT)
        {
          xs.resize(batchCount);
          KM_HIP(hipMemcpyAsync(xs.data(), pX, sizeof(void *) * batchCount,
                                hipMemcpyDeviceToHost, st));
        }
        KM_HIP(hipStreamSynchronize(st));
        std::vector<int> order(batchCount);
        std::iota(order.begin(), order.end(), 0);
        if (order_by == KM_ORDER_BY_INPUT)
        {
          // entries sharing an input become neighbours (the repeated X reads
          // hit L2); ties keep output order so equal outputs still merge on
          // chip
          std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            return xs[x] != xs[y] ? xs[x] < xs[y] : ys[x] < ys[y];
          });
        }
        else
          std::stable_sort(order.begin(), order.end(),
                           [&](int x, int y) { return ys[x] < ys[y]; });
        const long long target = (batchCount + G - 1) / G;
        std::vector<int> ranges;
        ranges.push_back(0);
        int noutputs = 0;
        if (order_by == KM_ORDER_BY_INPUT)
        {
          // outputs are shared between workgroups (atomic flush): even split
          std::vector<uintptr_t> u(ys);
          std::sort(u.begin(), u.end());
          noutputs = (int)(std::unique(u.begin(), u.end()) - u.begin());
          for (long long i = target; i < batchCount; i += target)
            ranges.push_back((int)i);
          ranges.push_back(batchCount);
        }
        else
        {
          // ranges end only where one output's run of entries ends
          std::vector<int> starts;
          std::vector<double> pre;
          for (int i = 0; i < batchCount;)
          {
            int j = i + 1;
            while (j < batchCount && ys[order[j]] == ys[order[i]])
              j++;
            starts.push_back(i);
            pre.push_back((double)i);
            i = j;
          }
          noutputs = (int)starts.size();
          starts.push_back(batchCount);
          pre.push_back((double)batchCount);
          balanced_cuts(starts, pre, G, ranges);
        }
        p->ngroups = (int)ranges.size() - 1;
        p->noutputs = noutputs;
        KM_HIP(hipMalloc(&p->order, sizeof(int) * batchCount));
        KM_HIP(hipMalloc(&p->ranges, sizeof(int) * ranges.size()));
        KM_HIP(hipMemcpy(p->order, order.data(), sizeof(int) * batchCount,
                         hipMemcpyHostToDevice));
        KM_HIP(hipMemcpy(p->ranges, ranges.data(), sizeof(int) * ranges.size(),
                         hipMemcpyHostToDevice));
      }
      if (p->owned)
      {
        // the binding: a device copy of the output pointers the plan was
        // built from, and the check kernel's flag
        KM_HIP(hipMalloc(&p->ybind, sizeof(void *) * batchCount));
        KM_HIP(hipMalloc(&p->flag, sizeof(int)));
        KM_HIP(hipMemcpyAsync(p->ybind, pY, sizeof(void *) * batchCount,
                              hipMemcpyDeviceToDevice, st));
        KM_HIP(hipMemsetAsync(p->flag, 0, sizeof(int), st));
        KM_HIP(hipStreamSynchronize(st));
      }
      return KM_SUCCESS;
    }();
    if (st2 != KM_SUCCESS)
    {
      km_plan_destroy(p);
      return st2;
    }
    *plan = p;
    return KM_SUCCESS;
  }

  km_status km_kronmult_pbatched_noalloc_f64(int d, int n, const double *A,
                                             double *const *pX,
                                             double *const *pY,
                                             double *const *pW, int batchCount,
                                             void *stream,
                                             const km_options *opt)
  {
    return pbatched_noalloc<double>(d, n, A, pX, pY, pW, batchCount, stream, opt);
  }

  km_status km_plan_destroy(km_plan plan)
  {
    if (!plan)
      return KM_SUCCESS;
    if (plan->order)
      (void)hipFree(plan->order);
    if (plan->ranges)
      (void)hipFree(plan->ranges);
    if (plan->ybind)
      (void)hipFree(plan->ybind);
    if (plan->flag)
      (void)hipFree(plan->flag);
    delete plan;
    return KM_SUCCESS;
  }

  km_status km_plan_info(km_plan plan, int *num_outputs, int *num_groups)
  {
    if (!plan)
      return KM_ERROR_INVALID_VALUE;
    if (num_outputs)
      *num_outputs = plan->noutputs;
    if (num_groups)
      *num_groups = plan->ngroups;
    return KM_SUCCESS;
  }

  km_status km_cooperative_faults(unsigned long long *count)
  {
    if (!count)
      return KM_ERROR_INVALID_VALUE;
    if (km_status g = km::runtime_check(); g != KM_SUCCESS)
      return g;
    KM_HIP(km::xcd_faults(count));
    return KM_SUCCESS;
  }

  km_status km_query_kernel(int d, int n, int elem_bytes, const km_options *opt,
                            int *family, int *threads)
  {
    return km_query_kernel_for(d, n, elem_bytes, 0, 0, opt, family, threads);
  }

  km_status km_query_kernel_for(int d, int n, int elem_bytes, int batchCount,
                                int owned, const km_options *opt, int *family,
                                int *threads)
  {
    km_status s = check_dn(d, n, 1);
    if (s != KM_SUCCESS || (elem_bytes != 4 && elem_bytes != 8) || batchCount < 0)
      return KM_ERROR_INVALID_VALUE;
    int m[6], nn[6];
    for (int k = 0; k < 6; k++)
      m[k] = nn[k] = n;
    Resolved r;
    if ((s = resolve(opt, d, m, nn, &r)) != KM_SUCCESS)
      return s;
    r.batch = batchCount;
    r.owned = owned ? 1 : 0;
    KernelInfo k, k2;
    bool onchip = (elem_bytes == 8) ? pick_onchip<double>(d, n, r, &k)
                                    : pick_onchip<float>(d, n, r, &k);
    bool twopass = false;
    if (!onchip && !r.force_global && !(r.fwd && r.exact))
      twopass = (elem_bytes == 8)
                    ? km::twopass_kernels<double>(d, n, r.exact, &k, &k2)
                    : km::twopass_kernels<float>(d, n, r.exact, &k, &k2);
    else if (!onchip && !r.force_global)
      twopass = (elem_bytes == 8) ? km::twopass_fwd_kernels<double>(d, n, &k, &k2)
                                  : km::twopass_fwd_kernels<float>(d, n, &k, &k2);
    if (twopass && !r.exact && xcd_device_ok() &&
        km::tuned_variant(d, n, elem_bytes, r.owned, batchCount) == KM_KERNEL_XCD)
    {
      KernelInfo kx;
      if ((elem_bytes == 8) ? km::xcd_kernel<double>(d, n, &kx)
                            : km::xcd_kernel<float>(d, n, &kx))
        k = kx;
    }
    if (!onchip && !twopass)
      k = (elem_bytes == 8) ? km::global_kernel<double>(r.exact)
                            : km::global_kernel<float>(r.exact);
    if (family)
      *family = k.family;
    if (threads)
      *threads = k.nt;
    return KM_SUCCESS;
  }

  km_status km_group_plan_create(km_group_plan *plan, int d, int batchCount,
                                 const int *n_host, const void *const *pY,
                                 int elem_bytes, void *stream)
  {
    if (!plan || d < 1 || d > 6 || batchCount < 0 || (batchCount > 0 && (!n_host || !pY)) ||
        (elem_bytes != 4 && elem_bytes != 8))
      return KM_ERROR_INVALID_VALUE;
    *plan = nullptr;
    if (km_status g = km::runtime_check(); g != KM_SUCCESS)
      return g;
    KernelInfo k;
    int nmax = 0, small_max = 0;
    bool ok = (elem_bytes == 8) ? km::runs_kernel<double>(d, false, &k, &nmax, &small_max)
                                : km::runs_kernel<float>(d, false, &k, &nmax, &small_max);
    if (!ok)
      return KM_ERROR_NOT_SUPPORTED;
    int pn = 0;
    for (int b = 0; b < batchCount; b++)
    {
      if (n_host[b] < 1)
        return KM_ERROR_INVALID_VALUE;
      if (n_host[b] > nmax)
        return KM_ERROR_NOT_SUPPORTED;
      pn = std::max(pn, n_host[b]);
    }
    km_group_plan p = new km_group_plan_s;
    p->d = d;
    p->batch = batchCount;
    p->elem_bytes = elem_bytes;
    p->nmax = pn;
    (void)hipGetDevice(&p->device);
    if (batchCount == 0)
    {
      *plan = p;
      return KM_SUCCESS;
    }
    hipStream_t st = (hipStream_t)stream;
    std::vector<uintptr_t> ys(batchCount);
    {
      km_status cs = [&]() -> km_status {
        KM_HIP(hipMemcpyAsync(ys.data(), pY, sizeof(void *) * batchCount,
                              hipMemcpyDeviceToHost, st));
        KM_HIP(hipStreamSynchronize(st));
        return KM_SUCCESS;
      }();
      if (cs != KM_SUCCESS)
      {
        delete p;
        return cs;
      }
    }
    // size bins, largest first (long entries are scheduled early), then
    // grouped by output; entry order kept inside an output (stable sort)
    std::vector<int> order(batchCount);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      if (n_host[x] != n_host[y])
        return n_host[x] > n_host[y];
      return ys[x] < ys[y];
    });
    // Time of an entry in the one-launch kernel (kron_runs), in units of one
    // byte at a workgroup's share of HBM bandwidth.  c = its X bytes plus its
    // flops at a quarter of a byte each.  Large entries have their X copied
    // in during the previous entry: c plus a small fixed term.  Small ones
    // run side by side on the workgroup's wavefronts but each waits on its
    // own loads and factor reads: (39,300 + 7 c) / waves, fitted to measured
    // d = 4 runs (n = 2 and n = 5 cost 0.39 and 1.1 us per entry where n = 8
    // costs 7.5).  Every run adds the fixed cost of its flush.  At d = 4 the
    // n = 8 entries take the register-blocked body (FMA calls) and weigh
    // 0.5 c, the double-buffered n = 6, 7 ones 1.2 c: swept on the 2..8 mix,
    // 0.202 ms with 1 and 1, 0.177 with 0.7 and 1, 0.151 with 0.5 and 1.2
    // (0.157-0.174 at the neighbouring points).
    const int waves = std::max(1, k.nt / 64);
    auto cost = [&](int b) {
      const double sz = std::pow((double)n_host[b], d);
      const double c = elem_bytes * sz + 0.5 * d * sz * n_host[b];
      if (waves > 1 && sz <= small_max)
        return (39300.0 + 7.0 * c) / waves;
      const double w = d != 4 ? 1.0 : (n_host[b] == 8 ? 0.5 : 1.2);
      return w * c + 2048.0;
    };
    constexpr double kRunCost = 25000.0;
    // A run of large entries costing more than kSplitShares workgroup shares
    // is cut into at most kSplitParts parts of about that size.  Measured on
    // d = 4, 20000 entries: 285 outputs of n = 8 only 530 -> 497 us (a
    // threshold of half a share: 397 us, but the 2..8 mix, whose runs are
    // shorter, 223 -> 240 us from the part stores and the combine launch).
    constexpr double kSplitShares = 2.0;
    constexpr int kSplitParts = 16;
    const long long G = std::max(
        1LL, std::min<long long>((long long)device_cus() * occupancy(k),
                                 batchCount));
    std::vector<int> ranges;
    ranges.push_back(0);
    int noutputs = 0;
    // runs: entries of one output and one size; then work units: a run of
    // large entries (whole-workgroup body) that would cost more than
    // kSplitShares workgroup shares is cut into parts that workgroups take
    // separately; a second launch adds their sums to the output in part
    // order
    std::vector<int> starts;    // unit starts (order positions)
    std::vector<int> sranges;   // ranges over the small entries' units
    std::vector<int> unitmeta;  // split id, part, parts
    std::vector<long long> splitoff;
    std::vector<int> splitfirst, splitparts;
    long long split_elems = 0;
    {
      std::vector<int> rstart;
      std::vector<double> rcost;
      double total = 0;
      for (int i = 0; i < batchCount;)
      {
        int j = i + 1;
        while (j < batchCount && ys[order[j]] == ys[order[i]] &&
               n_host[order[j]] == n_host[order[i]])
          j++;
        double c = kRunCost;
        for (int q = i; q < j; q++)
          c += cost(order[q]);
        rstart.push_back(i);
        rcost.push_back(c);
        total += c;
        noutputs++;
        i = j;
      }
      rstart.push_back(batchCount);
      const double share = total / (double)G;
      std::vector<double> pre;
      double acc = 0;
      for (size_t r = 0; r + 1 < rstart.size(); r++)
      {
        const int i = rstart[r], j = rstart[r + 1], k = j - i;
        const int n = n_host[order[i]];
        const double sz = std::pow((double)n, d);
        const bool large = !(waves > 1 && sz <= small_max);
        int parts = 1;
        if (large && k > 1 && rcost[r] > kSplitShares * share)
          parts = (int)std::min<double>(
              {(double)k, (double)kSplitParts, std::ceil(rcost[r] / (kSplitShares * share))});
        for (int q = 0; q < parts; q++)
        {
          const int us = i + (int)((long long)k * q / parts);
          const int ue = i + (int)((long long)k * (q + 1) / parts);
          starts.push_back(us);
          pre.push_back(acc);
          double c = kRunCost / parts;
          for (int e = us; e < ue; e++)
            c += cost(order[e]);
          acc += c;
          unitmeta.push_back(parts > 1 ? (int)splitoff.size() : -1);
          unitmeta.push_back(q);
          unitmeta.push_back(parts);
        }
        if (parts > 1)
        {
          splitoff.push_back(split_elems);
          splitfirst.push_back(i);
          splitparts.push_back(parts);
          split_elems += (long long)parts * (long long)sz;
        }
      }
      starts.push_back(batchCount);
      pre.push_back(acc);
      balanced_cuts(starts, pre, G, ranges);
      // the units of small entries (a suffix: sizes descend) on their own
      size_t us0 = 0;
      while (us0 + 1 < starts.size())
      {
        const double sz = std::pow((double)n_host[order[starts[us0]]], d);
        if (waves > 1 && sz <= small_max)
          break;
        ++us0;
      }
      if (us0 + 1 < starts.size())
      {
        std::vector<int> ss(starts.begin() + us0, starts.end());
        std::vector<double> sp(pre.begin() + us0, pre.end());
        for (double &v : sp)
          v -= pre[us0];
        sranges.push_back(ss[0]);
        balanced_cuts(ss, sp, G, sranges);
        for (int &c : sranges)
          c = (int)(std::lower_bound(starts.begin(), starts.end(), c) - starts.begin());
      }
      // the cuts fall on unit starts: express them in units
      for (int &c : ranges)
        c = (int)(std::lower_bound(starts.begin(), starts.end(), c) - starts.begin());
    }
    p->ngroups = (int)ranges.size() - 1;
    p->noutputs = noutputs;
    // per-size bins for the default multi-launch path: each bin split into
    // ranges of whole output runs, balanced inside the bin for the bin's
    // LDS-resident kernel
    std::vector<int> binranges;
    for (int i = 0; i < batchCount;)
    {
      const int n = n_host[order[i]];
      int j = i;
      while (j < batchCount && n_host[order[j]] == n)
        j++;
      KernelInfo kb;
      const bool have = (elem_bytes == 8) ? bin_kernel<double>(d, n, false, &kb)
                                          : bin_kernel<float>(d, n, false, &kb);
      const long long cnt = j - i;
      long long g = have ? (long long)device_cus() * occupancy(kb) : 1;
      // a bin that fits kWideWaves resident waves takes the wider kron_lds
      // for FMA calls, and its split is sized for it
      km_group_bin bn;
      bn.wide = 0;
      KernelInfo kw;
      if (have && kb.family == km::FAM_LDS && cnt <= kWideWaves * g &&
          ((elem_bytes == 8) ? km::lds_kernel_wide<double>(d, n, &kw)
                             : km::lds_kernel_wide<float>(d, n, &kw)))
      {
        bn.wide = 1;
        g = (long long)device_cus() * occupancy(kw);
      }
      g = std::max(1LL, std::min(g, cnt));
      bn.n = n;
      bn.start = i;
      bn.cnt = (int)cnt;
      bn.chunk = (int)((cnt + g - 1) / g);
      bn.first = (int)binranges.size();
      binranges.push_back(i);
      std::vector<int> starts;
      std::vector<double> pre;
      for (int q = i; q < j;)
      {
        int r = q + 1;
        while (r < j && ys[order[r]] == ys[order[q]])
          r++;
        starts.push_back(q);
        pre.push_back((double)(q - i));
        q = r;
      }
      starts.push_back(j);
      pre.push_back((double)(j - i));
      balanced_cuts(starts, pre, g, binranges);
      bn.count = (int)binranges.size() - bn.first - 1;
      p->bins.push_back(bn);
      // the next bin starts its own list (the shared boundary is repeated)
      i = j;
    }
    km_status st2 = [&]() -> km_status {
      KM_HIP(hipMalloc(&p->binranges, sizeof(int) * binranges.size()));
      KM_HIP(hipMemcpy(p->binranges, binranges.data(),
                       sizeof(int) * binranges.size(), hipMemcpyHostToDevice));
      KM_HIP(hipMalloc(&p->nvec, sizeof(int) * batchCount));
      KM_HIP(hipMalloc(&p->order, sizeof(int) * batchCount));
      KM_HIP(hipMalloc(&p->ranges, sizeof(int) * ranges.size()));
      KM_HIP(hipMalloc(&p->runstart, sizeof(int) * starts.size()));
      KM_HIP(hipMemcpy(p->runstart, starts.data(), sizeof(int) * starts.size(),
                       hipMemcpyHostToDevice));
      KM_HIP(hipMalloc(&p->unitmeta, sizeof(int) * unitmeta.size()));
      KM_HIP(hipMemcpy(p->unitmeta, unitmeta.data(), sizeof(int) * unitmeta.size(),
                       hipMemcpyHostToDevice));
      if (sranges.size() > 1)
      {
        KM_HIP(hipMalloc(&p->sranges, sizeof(int) * sranges.size()));
        KM_HIP(hipMemcpy(p->sranges, sranges.data(), sizeof(int) * sranges.size(),
                         hipMemcpyHostToDevice));
        p->nsgroups = (int)sranges.size() - 1;
        p->small_max = waves > 1 ? small_max : 0;
      }
      if (!splitoff.empty())
      {
        KM_HIP(hipMalloc(&p->splitws, (size_t)elem_bytes * (size_t)split_elems));
        KM_HIP(hipMalloc(&p->splitoff, sizeof(long long) * splitoff.size()));
        KM_HIP(hipMemcpy(p->splitoff, splitoff.data(), sizeof(long long) * splitoff.size(),
                         hipMemcpyHostToDevice));
        KM_HIP(hipMalloc(&p->splitfirst, sizeof(int) * splitfirst.size()));
        KM_HIP(hipMemcpy(p->splitfirst, splitfirst.data(), sizeof(int) * splitfirst.size(),
                         hipMemcpyHostToDevice));
        KM_HIP(hipMalloc(&p->splitparts, sizeof(int) * splitparts.size()));
        KM_HIP(hipMemcpy(p->splitparts, splitparts.data(), sizeof(int) * splitparts.size(),
                         hipMemcpyHostToDevice));
        p->nsplit = (int)splitoff.size();
      }
      KM_HIP(hipMemcpy(p->nvec, n_host, sizeof(int) * batchCount,
                       hipMemcpyHostToDevice));
      KM_HIP(hipMemcpy(p->order, order.data(), sizeof(int) * batchCount,
                       hipMemcpyHostToDevice));
      KM_HIP(hipMemcpy(p->ranges, ranges.data(), sizeof(int) * ranges.size(),
                       hipMemcpyHostToDevice));
      // the binding of the owned launch modes (as for km_plan_create)
      KM_HIP(hipMalloc(&p->ybind, sizeof(void *) * batchCount));
      KM_HIP(hipMalloc(&p->flag, sizeof(int)));
      KM_HIP(hipMemcpy(p->ybind, ys.data(), sizeof(void *) * batchCount,
                       hipMemcpyHostToDevice));
      KM_HIP(hipMemset(p->flag, 0, sizeof(int)));
      // different sizes never share an output, so the size bins may run
      // concurrently: up to four streams, bins dealt out largest first
      // (used by the owned-output path)
      const int ns = std::min<int>(4, (int)p->bins.size());
      if (ns > 1)
      {
        KM_HIP(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming));
        for (int i = 0; i < ns; i++)
        {
          hipStream_t st = nullptr;
          KM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
          p->streams.push_back(st);
          hipEvent_t ev = nullptr;
          KM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
          p->joins.push_back(ev);
        }
      }
      return KM_SUCCESS;
    }();
    if (st2 != KM_SUCCESS)
    {
      km_group_plan_destroy(p);
      return st2;
    }
    *plan = p;
    return KM_SUCCESS;
  }

  km_status km_group_plan_destroy(km_group_plan plan)
  {
    if (!plan)
      return KM_SUCCESS;
    if (plan->nvec)
      (void)hipFree(plan->nvec);
    if (plan->order)
      (void)hipFree(plan->order);
    if (plan->ranges)
      (void)hipFree(plan->ranges);
    if (plan->runstart)
      (void)hipFree(plan->runstart);
    if (plan->unitmeta)
      (void)hipFree(plan->unitmeta);
    if (plan->sranges)
      (void)hipFree(plan->sranges);
    if (plan->splitws)
      (void)hipFree(plan->splitws);
    if (plan->splitoff)
      (void)hipFree(plan->splitoff);
    if (plan->splitfirst)
      (void)hipFree(plan->splitfirst);
    if (plan->splitparts)
      (void)hipFree(plan->splitparts);
    if (plan->binranges)
      (void)hipFree(plan->binranges);
    if (plan->ybind)
      (void)hipFree(plan->ybind);
    if (plan->flag)
      (void)hipFree(plan->flag);
    for (hipStream_t st : plan->streams)
      (void)hipStreamDestroy(st);
    for (hipEvent_t ev : plan->joins)
      (void)hipEventDestroy(ev);
    if (plan->fork)
      (void)hipEventDestroy(plan->fork);
    delete plan;
    return KM_SUCCESS;
  }

  km_status km_group_plan_info(km_group_plan plan, int *num_outputs,
                               int *num_groups, int *max_n)
  {
    if (!plan)
      return KM_ERROR_INVALID_VALUE;
    if (num_outputs)
      *num_outputs = plan->noutputs;
    if (num_groups)
      *num_groups = plan->ngroups;
    if (max_n)
      *max_n = plan->nmax;
    return KM_SUCCESS;
  }

  km_status km_kronmult_grouped_f64(km_group_plan plan,
                                    const double *const *Aarray,
                                    double *const *pX, double *const *pY,
                                    void *stream, const km_options *opt)
  {
    return grouped<double>(plan, Aarray, pX, pY, stream, opt);
  }

  km_status km_kronmult_grouped_f32(km_group_plan plan,
                                    const float *const *Aarray, float *const *pX,
                                    float *const *pY, void *stream,
                                    const km_options *opt)
  {
    return grouped<float>(plan, Aarray, pX, pY, stream, opt);
  }

  km_status km_kgemm_nn_batched_f64(int mm, int nn, int kk, double alpha,
                                     const double *const *Aarray,
                                     const int *ldAarray,
                                     const double *const *Barray,
                                     const int *ldBarray, double beta,
                                     double *const *Carray, const int *ldCarray,
                                     int batchCount, void *stream,
                                     const km_options *opt)
  {
    return kgemm<double>(0, mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray,
                      beta, Carray, ldCarray, batchCount, stream, opt);
  }

  km_status km_kgemm_nn_batched_f32(int mm, int nn, int kk, float alpha,
                                     const float *const *Aarray,
                                     const int *ldAarray,
                                     const float *const *Barray,
                                     const int *ldBarray, float beta,
                                     float *const *Carray, const int *ldCarray,
                                     int batchCount, void *stream,
                                     const km_options *opt)
  {
    return kgemm<float>(0, mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray,
                      beta, Carray, ldCarray, batchCount, stream, opt);
  }

  km_status km_kgemm_nt_batched_f64(int mm, int nn, int kk, double alpha,
                                     const double *const *Aarray,
                                     const int *ldAarray,
                                     const double *const *Barray,
                                     const int *ldBarray, double beta,
                                     double *const *Carray, const int *ldCarray,
                                     int batchCount, void *stream,
                                     const km_options *opt)
  {
    return kgemm<double>(1, mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray,
                      beta, Carray, ldCarray, batchCount, stream, opt);
  }

  km_status km_kgemm_nt_batched_f32(int mm, int nn, int kk, float alpha,
                                     const float *const *Aarray,
                                     const int *ldAarray,
                                     const float *const *Barray,
                                     const int *ldBarray, float beta,
                                     float *const *Carray, const int *ldCarray,
                                     int batchCount, void *stream,
                                     const km_options *opt)
  {
    return kgemm<float>(1, mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray,
                      beta, Carray, ldCarray, batchCount, stream, opt);
  }

  km_status km_kgemm_tt_batched_f64(int mm, int nn, int kk, double alpha,
                                     const double *const *Aarray,
                                     const int *ldAarray,
                                     const double *const *Barray,
                                     const int *ldBarray, double beta,
                                     double *const *Carray, const int *ldCarray,
                                     int batchCount, void *stream,
                                     const km_options *opt)
  {
    return kgemm<double>(2, mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray,
                      beta, Carray, ldCarray, batchCount, stream, opt);
  }

  km_status km_kgemm_tt_batched_f32(int mm, int nn, int kk, float alpha,
                                     const float *const *Aarray,
                                     const int *ldAarray,
                                     const float *const *Barray,
                                     const int *ldBarray, float beta,
                                     float *const *Carray, const int *ldCarray,
                                     int batchCount, void *stream,
                                     const km_options *opt)
  {
    return kgemm<float>(2, mm, nn, kk, alpha, Aarray, ldAarray, Barray, ldBarray,
                      beta, Carray, ldCarray, batchCount, stream, opt);
  }

  km_status km_xbatched_tables_f64(
      int d, int batchCount, const double *const *Apool, const long long *a_stride,
      const int *a_count, const int *aidx, double *X, long long x_stride,
      int x_count, const int *xidx, double *Y, long long y_stride, int y_count,
      const int *yidx, const double **Aarray, double **pX, double **pY, int *bad,
      void *stream)
  {
    return xbatched_tables<double>(d, batchCount, Apool, a_stride, a_count, aidx, X,
                               x_stride, x_count, xidx, Y, y_stride, y_count,
                               yidx, Aarray, pX, pY, bad, stream);
  }

  km_status km_xbatched_tables_f32(
      int d, int batchCount, const float *const *Apool, const long long *a_stride,
      const int *a_count, const int *aidx, float *X, long long x_stride,
      int x_count, const int *xidx, float *Y, long long y_stride, int y_count,
      const int *yidx, const float **Aarray, float **pX, float **pY, int *bad,
      void *stream)
  {
    return xbatched_tables<float>(d, batchCount, Apool, a_stride, a_count, aidx, X,
                               x_stride, x_count, xidx, Y, y_stride, y_count,
                               yidx, Aarray, pX, pY, bad, stream);
  }

  double km_kron_flops(int d, const int *m, const int *n)
  {
    // kroncommon.hpp:154-185: stage 1 applies A_1 to n_2..n_d vectors of
    // length n_1, then m_1 copies of the (d-1)-dimensional problem follow
    if (d < 1 || !m || !n)
      return 0;
    if (d == 1)
      return 2.0 * m[0] * n[0];
    double rest = 1;
    for (int k = 1; k < d; k++)
      rest *= n[k];
    return 2.0 * rest * n[0] * m[0] + km_kron_flops(d - 1, m + 1, n + 1) * m[0];
  }

} // extern "C"
