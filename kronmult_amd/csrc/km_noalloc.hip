// km_noalloc.hip -- the static per-device buffer of the allocation-free
// entry point (km_kronmult_pbatched_noalloc_f64).
//
// The buffer holds the device-built output order of one call (km_order.hip):
// sort keys and values, run starts, histograms and workgroup ranges for up to
// KM_NOALLOC_MAX_BATCH entries.  Each GPU the library runs on holds its own
// copy until the process ends.  Every call on a GPU shares that GPU's copy,
// so calls that overlap on one GPU race; the order builder writes every word
// it reads, so nothing needs clearing between calls.
#include <map>
#include <mutex>

#include "km_internal.h"

namespace km
{
namespace
{

// Host code never passes this name to a kernel: on the host it is only a
// stand-in symbol.  The device address comes from hipGetSymbolAddress for the
// current device (noalloc_scratch below).
__device__ __align__(256) unsigned char scratch[KM_NOALLOC_SCRATCH_BYTES];

} // namespace

hipError_t noalloc_scratch(void **ptr)
{
  static std::mutex mu;
  static std::map<int, void *> addr; // device -> address of its copy
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return e;
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(dev);
  if (it == addr.end())
  {
    // first call on this device, with the device current
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return e;
    it = addr.emplace(dev, p).first;
  }
  *ptr = it->second;
  return hipSuccess;
}

} // namespace km
