// kron_reg instantiations (register-blocked mode groups, n a power of two).
#include "km_internal.h"

namespace km
{

#define KM_REG_CASE(T, D, N, RM)                                               \
  if (d == D && n == N)                                                        \
  {                                                                            \
    out->fn = exact ? (const void *)&kron_reg<T, D, N, RM, true>               \
                    : (const void *)&kron_reg<T, D, N, RM, false>;             \
    out->nt = cpow(N, D - RM);                                                 \
    out->family = FAM_REG;                                                     \
    return true;                                                               \
  }

template<>
bool reg_kernel<double>(int d, int n, bool exact, KernelInfo *out)
{
  KM_REG_CASE(double, 6, 4, 2)
  KM_REG_CASE(double, 3, 8, 1)
  KM_REG_CASE(double, 4, 8, 1)
  KM_REG_CASE(double, 4, 4, 1)
  return false;
}

template<>
bool reg_kernel<float>(int d, int n, bool exact, KernelInfo *out)
{
  KM_REG_CASE(float, 6, 4, 2)
  KM_REG_CASE(float, 3, 8, 1)
  KM_REG_CASE(float, 4, 8, 1)
  KM_REG_CASE(float, 4, 4, 1)
  return false;
}

} // namespace km
