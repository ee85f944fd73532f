// kron_slab instantiations: shapes whose n^d tensor exceeds the LDS budget
// while one slab of n^(d-1) elements (double buffered) still fits.
#include "km_internal.h"

namespace km
{

#define KM_SLAB_CASE(T, D, N, NT)                                              \
  if (d == D && n == N)                                                        \
  {                                                                            \
    out->fn = exact ? (const void *)&kron_slab<T, D, N, NT, true>              \
                    : (const void *)&kron_slab<T, D, N, NT, false>;            \
    out->nt = NT;                                                              \
    out->family = FAM_SLAB;                                                    \
    return true;                                                               \
  }

template<>
bool slab_kernel<double>(int d, int n, bool exact, KernelInfo *out)
{
  KM_SLAB_CASE(double, 6, 6, 512)
  KM_SLAB_CASE(double, 5, 8, 512)
  return false;
}

template<>
bool slab_kernel<float>(int d, int n, bool exact, KernelInfo *out)
{
  KM_SLAB_CASE(float, 6, 6, 512)
  KM_SLAB_CASE(float, 5, 8, 512)
  return false;
}

#define KM_SLAB_FWD_CASE(T, D, N, NT)                                          \
  if (d == D && n == N)                                                        \
  {                                                                            \
    out->fn = (const void *)&kron_slab_fwd<T, D, N, NT>;                       \
    out->nt = NT;                                                              \
    out->family = FAM_SLAB;                                                    \
    return true;                                                               \
  }

template<>
bool slab_fwd_kernel<double>(int d, int n, KernelInfo *out)
{
  KM_SLAB_FWD_CASE(double, 6, 6, 512)
  KM_SLAB_FWD_CASE(double, 5, 8, 512)
  return false;
}

template<>
bool slab_fwd_kernel<float>(int d, int n, KernelInfo *out)
{
  KM_SLAB_FWD_CASE(float, 6, 6, 512)
  KM_SLAB_FWD_CASE(float, 5, 8, 512)
  return false;
}

} // namespace km
