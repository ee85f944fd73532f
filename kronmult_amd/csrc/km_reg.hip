// kron_reg instantiations (register-blocked mode groups).  d4n6 (216
// threads) is a default; d4n5 and d4n7 (125 and 343 threads) only serve
// owned grouped bins and tuned variants: plain calls measured slower there
// than kron_lds (d4n5 x65536 4:1 0.134 against 0.121 ms, d4n7 x20000 0.214
// against 0.173 ms), owned bins of 2857 entries faster (d4n5 0.018 against
// 0.032 ms, d4n7 0.057 against 0.073 ms); d4n6 x65536 0.310 against 0.341.
#include "km_internal.h"

namespace km
{

// CAP: resident workgroups per CU the grid is sized for (0: all that fit).
// d3n8 (61 VGPRs, 64 threads: 32 fit) measured best at 24: x262144 4:1
// 0.461 ms at 24 (where its registers held it before the transposes used
// one address per thread), 0.523 ms at 32.
#define KM_REG_CASE(T, D, N, RM, CAP)                                          \
  if (d == D && n == N)                                                        \
  {                                                                            \
    out->fn = exact ? (const void *)&kron_reg<T, D, N, RM, true>               \
                    : (const void *)&kron_reg<T, D, N, RM, false>;             \
    out->nt = cpow(N, D - RM);                                                 \
    out->family = FAM_REG;                                                     \
    out->max_per_cu = CAP;                                                     \
    return true;                                                               \
  }

template<>
bool reg_kernel<double>(int d, int n, bool exact, KernelInfo *out, bool all)
{
  KM_REG_CASE(double, 6, 4, 2, 0)
  KM_REG_CASE(double, 3, 8, 1, 24)
  KM_REG_CASE(double, 4, 8, 1, 0)
  KM_REG_CASE(double, 4, 4, 1, 0)
  KM_REG_CASE(double, 4, 6, 1, 0)
  if (all)
  {
    KM_REG_CASE(double, 4, 5, 1, 0)
    KM_REG_CASE(double, 4, 7, 1, 0)
  }
  return false;
}

template<>
bool reg_kernel<float>(int d, int n, bool exact, KernelInfo *out, bool all)
{
  KM_REG_CASE(float, 6, 4, 2, 0)
  KM_REG_CASE(float, 3, 8, 1, 24)
  KM_REG_CASE(float, 4, 8, 1, 0)
  KM_REG_CASE(float, 4, 4, 1, 0)
  KM_REG_CASE(float, 4, 6, 1, 0)
  if (all)
  {
    KM_REG_CASE(float, 4, 5, 1, 0)
    KM_REG_CASE(float, 4, 7, 1, 0)
  }
  return false;
}

} // namespace km
