// km_tune.cpp -- the kernel-variant table and its autotuner.
//
// Several kernels can serve one square (d, n) FMA call: the slab-streamed
// kron_slab or the split-pair kron_half (one or two slabs per step) above the
// LDS size; kron_reg, kron_lds or the wider kron_lds below it.  Which is
// fastest depends on the batch size and on the reduction (plan-owned outputs
// or atomics).  The table maps (d, n, element size, owned, batch bin =
// floor(log2 batch)) to a variant; pick_onchip consults it for FMA calls and
// falls back to the built-in rule where it has no entry.  km_tune_shape fills
// an entry by timing every candidate on synthetic device data through the
// public entry point; the table is saved to and loaded from a JSON file keyed
// by the GPU architecture (KRONMULT_AMD_TUNE_FILE is loaded on first use).
// No counterpart in the reference, whose launch geometry is fixed
// (kronmult6_batched.cpp:8-14, kronmult6_xbatched.cpp:12-23).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "km_internal.h"

namespace km
{
namespace
{

struct Key
{
  int d, n, elem, owned, bin;
  bool operator<(const Key &o) const
  {
    return std::tie(d, n, elem, owned, bin) < std::tie(o.d, o.n, o.elem, o.owned, o.bin);
  }
};

struct Entry
{
  int variant = KM_KERNEL_DEFAULT;
  std::vector<std::pair<int, double>> ms; // measured candidates (ms per call)
};

std::mutex g_mu;
std::map<Key, Entry> g_table;
bool g_env_loaded = false;
// variant forced on this thread while the tuner measures a candidate
thread_local int g_force = KM_KERNEL_DEFAULT;

const char *kNames[] = {"default", "slab",     "half1", "half2", "reg",
                        "lds",     "lds_wide", "twopass", "xcd"};
constexpr int kNumVariants = sizeof(kNames) / sizeof(kNames[0]);

int variant_of(const char *s)
{
  for (int i = 0; i < kNumVariants; i++)
    if (std::strcmp(s, kNames[i]) == 0)
      return i;
  return -1;
}

std::string arch_name()
{
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return "unknown";
  std::string a = prop.gcnArchName;
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}

km_status load_file(const char *path, bool replace);

void load_env_locked()
{
  if (g_env_loaded)
    return;
  g_env_loaded = true;
  const char *p = std::getenv("KRONMULT_AMD_TUNE_FILE");
  if (p && *p)
  {
    g_mu.unlock();
    (void)load_file(p, false);
    g_mu.lock();
  }
}

// Minimal reader of the files km_tune_save writes: one entry per line.
km_status load_file(const char *path, bool replace)
{
  FILE *f = std::fopen(path, "r");
  if (!f)
    return KM_ERROR_INVALID_VALUE;
  std::string text;
  char buf[4096];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0)
    text.append(buf, got);
  std::fclose(f);
  const size_t ap = text.find("\"arch\"");
  if (ap == std::string::npos)
    return KM_ERROR_INVALID_VALUE;
  char arch[64] = "";
  if (std::sscanf(text.c_str() + ap, "\"arch\": \"%63[^\"]\"", arch) != 1)
    return KM_ERROR_INVALID_VALUE;
  if (arch_name() != arch)
    return KM_ERROR_NOT_SUPPORTED; // tuned on another GPU
  std::map<Key, Entry> add;
  size_t pos = 0;
  while ((pos = text.find("{\"d\":", pos)) != std::string::npos)
  {
    Key k{};
    char name[32] = "";
    if (std::sscanf(text.c_str() + pos,
                    "{\"d\": %d, \"n\": %d, \"elem\": %d, \"owned\": %d, \"bin\": %d, "
                    "\"variant\": \"%31[^\"]\"",
                    &k.d, &k.n, &k.elem, &k.owned, &k.bin, name) == 6)
    {
      const int v = variant_of(name);
      if (v >= 0)
        add[k].variant = v;
    }
    pos += 5;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (replace)
    g_table.clear();
  for (auto &e : add)
    g_table[e.first] = e.second;
  return KM_SUCCESS;
}

} // namespace

int batch_bin(long long batch)
{
  return batch <= 1 ? 0 : 63 - __builtin_clzll((unsigned long long)batch);
}

int tuned_variant(int d, int n, int elem_bytes, int owned, long long batch)
{
  if (g_force != KM_KERNEL_DEFAULT)
    return g_force;
  std::lock_guard<std::mutex> lk(g_mu);
  load_env_locked();
  auto it = g_table.find(Key{d, n, elem_bytes, owned ? 1 : 0, batch_bin(batch)});
  return it == g_table.end() ? KM_KERNEL_DEFAULT : it->second.variant;
}

// Variants that can serve an FMA call of square (d, n).
template<typename T>
std::vector<int> candidates(int d, int n)
{
  std::vector<int> c;
  KernelInfo k;
  if (xcd_kernel<T>(d, n, &k))
  {
    c = {KM_KERNEL_XCD, KM_KERNEL_TWOPASS};
    return c;
  }
  if (half_kernel<T>(d, n, 1, &k))
  {
    c = {KM_KERNEL_SLAB, KM_KERNEL_HALF1, KM_KERNEL_HALF2};
    return c;
  }
  if (reg_kernel<T>(d, n, false, &k))
    c.push_back(KM_KERNEL_REG);
  if (lds_kernel<T>(d, n, false, false, &k))
    c.push_back(KM_KERNEL_LDS);
  if (lds_kernel_wide<T>(d, n, &k))
    c.push_back(KM_KERNEL_LDS_WIDE);
  return c;
}

template<typename T>
bool variant_kernel(int v, int d, int n, KernelInfo *k)
{
  switch (v)
  {
  case KM_KERNEL_SLAB: return slab_kernel<T>(d, n, false, k);
  case KM_KERNEL_HALF1: return half_kernel<T>(d, n, 1, k);
  case KM_KERNEL_HALF2: return half_kernel<T>(d, n, 2, k);
  case KM_KERNEL_REG: return reg_kernel<T>(d, n, false, k, true);
  case KM_KERNEL_LDS: return lds_kernel<T>(d, n, false, false, k);
  case KM_KERNEL_LDS_WIDE: return lds_kernel_wide<T>(d, n, k);
  case KM_KERNEL_TWOPASS:
  {
    KernelInfo k2;
    return twopass_kernels<T>(d, n, false, k, &k2);
  }
  case KM_KERNEL_XCD: return xcd_kernel<T>(d, n, k);
  default: return false;
  }
}
template bool variant_kernel<double>(int, int, int, KernelInfo *);
template bool variant_kernel<float>(int, int, int, KernelInfo *);

} // namespace km

using namespace km;

namespace
{

// fill n elements with a fixed pattern (finite values of modest size)
km_status fill(void *p, size_t bytes, hipStream_t st)
{
  // 0x3F800000: 1.0f as float, about 0.0078 as the high word of a double
  return hipMemsetD32Async((hipDeviceptr_t)p, 0x3F800000, bytes / 4, st) == hipSuccess
             ? KM_SUCCESS
             : KM_ERROR_HIP;
}

template<typename T>
km_status measure(int d, int n, int batch, int overlap, int owned, int v,
                  hipStream_t st, double *ms)
{
  long long s = 1;
  for (int i = 0; i < d; i++)
    s *= n;
  const int nY = std::max(1, batch / std::max(1, overlap));
  T *A = nullptr, *X = nullptr, *Y = nullptr, *W = nullptr;
  T **pX = nullptr, **pY = nullptr, **pW = nullptr;
  // the two-pass shapes take one n^d workspace per entry
  KernelInfo kw;
  const bool need_w = xcd_kernel<T>(d, n, &kw);
  km_plan plan = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  km_status st2 = KM_SUCCESS;
  auto hip = [&](hipError_t e) {
    if (e != hipSuccess && st2 == KM_SUCCESS)
      st2 = KM_ERROR_HIP;
    return e == hipSuccess;
  };
  const size_t abytes = sizeof(T) * (size_t)batch * d * n * n;
  const size_t xbytes = sizeof(T) * (size_t)batch * s;
  const size_t ybytes = sizeof(T) * (size_t)nY * s;
  if (hip(hipMalloc(&A, abytes)) && hip(hipMalloc(&X, xbytes)) && hip(hipMalloc(&Y, ybytes)) &&
      hip(hipMalloc(&pX, sizeof(T *) * batch)) && hip(hipMalloc(&pY, sizeof(T *) * batch)) &&
      (!need_w || (hip(hipMalloc(&W, xbytes)) && hip(hipMalloc(&pW, sizeof(T *) * batch)))))
  {
    std::vector<T *> hx(batch), hy(batch), hw(batch);
    for (int b = 0; b < batch; b++)
    {
      hx[b] = X + (size_t)b * s;
      hy[b] = Y + (size_t)(b / std::max(1, overlap) % nY) * s;
      hw[b] = need_w ? W + (size_t)b * s : nullptr;
    }
    hip(hipMemcpy(pX, hx.data(), sizeof(T *) * batch, hipMemcpyHostToDevice));
    hip(hipMemcpy(pY, hy.data(), sizeof(T *) * batch, hipMemcpyHostToDevice));
    if (need_w)
      hip(hipMemcpy(pW, hw.data(), sizeof(T *) * batch, hipMemcpyHostToDevice));
    if (st2 == KM_SUCCESS)
      st2 = fill(A, abytes, st);
    if (st2 == KM_SUCCESS)
      st2 = fill(X, xbytes, st);
    if (st2 == KM_SUCCESS)
      st2 = fill(Y, ybytes, st);
    g_force = v;
    if (st2 == KM_SUCCESS && owned)
      st2 = km_plan_create(&plan, d, n, sizeof(T), (const void *const *)pY, batch, st);
    km_options o;
    km_options_init(&o);
    o.plan = plan;
    auto call = [&]() {
      return (sizeof(T) == 8)
                 ? km_kronmult_pbatched_f64(d, n, (const double *)A, (double *const *)pX,
                                            (double *const *)pY, (double *const *)pW, batch,
                                            st, &o)
                 : km_kronmult_pbatched_f32(d, n, (const float *)A, (float *const *)pX,
                                            (float *const *)pY, (float *const *)pW, batch, st,
                                            &o);
    };
    for (int w = 0; w < 3 && st2 == KM_SUCCESS; w++)
      st2 = call();
    const int reps = 10;
    if (st2 == KM_SUCCESS && hip(hipEventCreate(&e0)) && hip(hipEventCreate(&e1)) &&
        hip(hipEventRecord(e0, st)))
    {
      for (int r = 0; r < reps && st2 == KM_SUCCESS; r++)
        st2 = call();
      float t = 0;
      if (st2 == KM_SUCCESS && hip(hipEventRecord(e1, st)) && hip(hipEventSynchronize(e1)) &&
          hip(hipEventElapsedTime(&t, e0, e1)))
        *ms = t / reps;
    }
    g_force = KM_KERNEL_DEFAULT;
  }
  if (plan)
    km_plan_destroy(plan);
  if (e0)
    (void)hipEventDestroy(e0);
  if (e1)
    (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(st);
  (void)hipFree(A);
  (void)hipFree(X);
  (void)hipFree(Y);
  (void)hipFree(pX);
  (void)hipFree(pY);
  (void)hipFree(W);
  (void)hipFree(pW);
  return st2;
}

template<typename T>
km_status tune(int d, int n, int batch, int overlap, int owned, hipStream_t st, int *out)
{
  const std::vector<int> cand = candidates<T>(d, n);
  if (cand.empty())
    return KM_ERROR_NOT_SUPPORTED;
  Entry e;
  double best = 0;
  for (int v : cand)
  {
    double ms = 0;
    km_status s = measure<T>(d, n, batch, overlap, owned, v, st, &ms);
    if (s != KM_SUCCESS)
      return s;
    e.ms.push_back({v, ms});
    if (e.variant == KM_KERNEL_DEFAULT || ms < best)
    {
      e.variant = v;
      best = ms;
    }
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    load_env_locked();
    g_table[Key{d, n, (int)sizeof(T), owned ? 1 : 0, batch_bin(batch)}] = e;
  }
  if (out)
    *out = e.variant;
  return KM_SUCCESS;
}

} // namespace

extern "C"
{

  km_status km_tune_shape(int d, int n, int elem_bytes, int batchCount,
                          int overlap, int owned, void *stream, int *variant)
  {
    if (d < 1 || d > 6 || n < 1 || batchCount < 1 || overlap < 1 ||
        (elem_bytes != 4 && elem_bytes != 8))
      return KM_ERROR_INVALID_VALUE;
    if (km_status g = runtime_check(); g != KM_SUCCESS)
      return g;
    return elem_bytes == 8
               ? tune<double>(d, n, batchCount, overlap, owned, (hipStream_t)stream, variant)
               : tune<float>(d, n, batchCount, overlap, owned, (hipStream_t)stream, variant);
  }

  km_status km_tune_set(int d, int n, int elem_bytes, int batchCount, int owned,
                        int variant)
  {
    if (d < 1 || d > 6 || n < 1 || batchCount < 1 || (elem_bytes != 4 && elem_bytes != 8) ||
        variant < 0 || variant >= kNumVariants)
      return KM_ERROR_INVALID_VALUE;
    KernelInfo k;
    if (variant != KM_KERNEL_DEFAULT &&
        !(elem_bytes == 8 ? variant_kernel<double>(variant, d, n, &k)
                          : variant_kernel<float>(variant, d, n, &k)))
      return KM_ERROR_NOT_SUPPORTED;
    std::lock_guard<std::mutex> lk(g_mu);
    load_env_locked();
    const Key key{d, n, elem_bytes, owned ? 1 : 0, batch_bin(batchCount)};
    if (variant == KM_KERNEL_DEFAULT)
      g_table.erase(key);
    else
      g_table[key] = Entry{variant, {}};
    return KM_SUCCESS;
  }

  km_status km_tune_get(int d, int n, int elem_bytes, int batchCount, int owned,
                        int *variant)
  {
    if (!variant || batchCount < 1)
      return KM_ERROR_INVALID_VALUE;
    *variant = tuned_variant(d, n, elem_bytes, owned, batchCount);
    return KM_SUCCESS;
  }

  km_status km_tune_clear(void)
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_env_loaded = true;
    g_table.clear();
    return KM_SUCCESS;
  }

  km_status km_tune_save(const char *path)
  {
    if (!path)
      return KM_ERROR_INVALID_VALUE;
    FILE *f = std::fopen(path, "w");
    if (!f)
      return KM_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> lk(g_mu);
    std::fprintf(f, "{\n  \"arch\": \"%s\",\n  \"entries\": [\n", arch_name().c_str());
    size_t i = 0;
    for (const auto &e : g_table)
    {
      std::fprintf(f,
                   "    {\"d\": %d, \"n\": %d, \"elem\": %d, \"owned\": %d, \"bin\": %d, "
                   "\"variant\": \"%s\", \"ms\": {",
                   e.first.d, e.first.n, e.first.elem, e.first.owned, e.first.bin,
                   kNames[e.second.variant]);
      for (size_t j = 0; j < e.second.ms.size(); j++)
        std::fprintf(f, "%s\"%s\": %.5f", j ? ", " : "", kNames[e.second.ms[j].first],
                     e.second.ms[j].second);
      std::fprintf(f, "}}%s\n", ++i < g_table.size() ? "," : "");
    }
    std::fprintf(f, "  ]\n}\n");
    std::fclose(f);
    return KM_SUCCESS;
  }

  km_status km_tune_load(const char *path)
  {
    if (!path)
      return KM_ERROR_INVALID_VALUE;
    return load_file(path, true);
  }

} // extern "C"
