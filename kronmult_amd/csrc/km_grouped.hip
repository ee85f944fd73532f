// kron_grouped instantiations: per-entry sizes n_b <= NMAX for each d, where
// NMAX^d fits the LDS budget.
#include "km_internal.h"

namespace km
{

#define KM_GROUPED_CASE(T, D, NMAX, NT)                                        \
  if (d == D)                                                                  \
  {                                                                            \
    out->fn = exact ? (const void *)&kron_grouped<T, D, NMAX, NT, true>        \
                    : (const void *)&kron_grouped<T, D, NMAX, NT, false>;      \
    out->nt = NT;                                                              \
    out->family = FAM_GROUPED;                                                 \
    *nmax = NMAX;                                                              \
    return true;                                                               \
  }

template<>
bool grouped_kernel<double>(int d, bool exact, KernelInfo *out, int *nmax)
{
  KM_GROUPED_CASE(double, 1, 8, 64)
  KM_GROUPED_CASE(double, 2, 8, 64)
  KM_GROUPED_CASE(double, 3, 8, 64)
  KM_GROUPED_CASE(double, 4, 8, 256)
  KM_GROUPED_CASE(double, 5, 6, 256)
  KM_GROUPED_CASE(double, 6, 4, 256)
  return false;
}

template<>
bool grouped_kernel<float>(int d, bool exact, KernelInfo *out, int *nmax)
{
  KM_GROUPED_CASE(float, 1, 8, 64)
  KM_GROUPED_CASE(float, 2, 8, 64)
  KM_GROUPED_CASE(float, 3, 8, 64)
  KM_GROUPED_CASE(float, 4, 8, 256)
  KM_GROUPED_CASE(float, 5, 6, 256)
  KM_GROUPED_CASE(float, 6, 4, 256)
  return false;
}

} // namespace km
