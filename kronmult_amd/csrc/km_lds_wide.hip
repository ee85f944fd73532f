// kron_lds instantiations with wider workgroups (lds_threads_wide) for the
// grouped call's size bins whose entries fit a few resident waves of
// workgroups: there one entry's chain of loads and passes bounds the launch,
// and more threads per entry shorten it.  FMA, backward mode order only; the
// (d, n) listed are those where the width differs from lds_threads and
// kron_lds (not kron_reg) serves the size.
#include "km_internal.h"

namespace km
{

#define KM_LDS_WIDE_CASE(T, D, N)                                              \
  if (d == D && n == N)                                                        \
  {                                                                            \
    constexpr int NT_ = lds_threads_wide(cpow(N, D));                          \
    static_assert(NT_ != lds_threads(cpow(N, D)), "same width as kron_lds");   \
    out->fn = (const void *)&kron_lds<T, D, N, NT_, false, false>;             \
    out->nt = NT_;                                                             \
    out->family = FAM_LDS;                                                     \
    return true;                                                               \
  }

template<>
bool lds_kernel_wide<double>(int d, int n, KernelInfo *out)
{
  KM_LDS_WIDE_CASE(double, 3, 7)
  KM_LDS_WIDE_CASE(double, 4, 5)
  KM_LDS_WIDE_CASE(double, 4, 7)
  KM_LDS_WIDE_CASE(double, 5, 4)
  KM_LDS_WIDE_CASE(double, 5, 5)
  KM_LDS_WIDE_CASE(double, 6, 3)
  return false;
}

template<>
bool lds_kernel_wide<float>(int d, int n, KernelInfo *out)
{
  KM_LDS_WIDE_CASE(float, 3, 7)
  KM_LDS_WIDE_CASE(float, 4, 5)
  KM_LDS_WIDE_CASE(float, 4, 7)
  KM_LDS_WIDE_CASE(float, 5, 4)
  KM_LDS_WIDE_CASE(float, 5, 5)
  KM_LDS_WIDE_CASE(float, 6, 3)
  return false;
}

} // namespace km
