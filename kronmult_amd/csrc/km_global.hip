// kron_global instantiations (any shape; stages through caller workspace) and
// the d -> per-d unit dispatch of kron_lds.
#include "km_internal.h"

namespace km
{

template<>
KernelInfo global_kernel<double>(bool exact)
{
  KernelInfo k;
  k.fn = exact ? (const void *)&kron_global<double, true>
               : (const void *)&kron_global<double, false>;
  k.nt = 256;
  k.family = FAM_GLOBAL;
  return k;
}

template<>
KernelInfo global_kernel<float>(bool exact)
{
  KernelInfo k;
  k.fn = exact ? (const void *)&kron_global<float, true>
               : (const void *)&kron_global<float, false>;
  k.nt = 256;
  k.family = FAM_GLOBAL;
  return k;
}

template<typename T>
bool lds_kernel(int d, int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (d)
  {
  case 1: return lds_kernel_d<T, 1>(n, exact, fwd, out);
  case 2: return lds_kernel_d<T, 2>(n, exact, fwd, out);
  case 3: return lds_kernel_d<T, 3>(n, exact, fwd, out);
  case 4: return lds_kernel_d<T, 4>(n, exact, fwd, out);
  case 5: return lds_kernel_d<T, 5>(n, exact, fwd, out);
  case 6: return lds_kernel_d<T, 6>(n, exact, fwd, out);
  }
  return false;
}

template bool lds_kernel<double>(int, int, bool, bool, KernelInfo *);
template bool lds_kernel<float>(int, int, bool, bool, KernelInfo *);

} // namespace km
