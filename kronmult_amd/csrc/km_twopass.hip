// Two-pass family instantiations (kron_sub + kron_pair): square shapes whose
// n^(d-1) slab does not fit LDS while n^(d-2) does.
#include "km_internal.h"

namespace km
{

#define KM_TWOPASS_CASE(T, D, N, NT1, NT2)                                     \
  if (d == D && n == N)                                                        \
  {                                                                            \
    sub->fn = exact ? (const void *)&kron_sub<T, D, N, NT1, true>              \
                    : (const void *)&kron_sub<T, D, N, NT1, false>;            \
    sub->nt = NT1;                                                             \
    sub->family = FAM_TWOPASS;                                                 \
    pair->fn = exact ? (const void *)&kron_pair<T, D, N, NT2, true>            \
                     : (const void *)&kron_pair<T, D, N, NT2, false>;          \
    pair->nt = NT2;                                                            \
    pair->family = FAM_TWOPASS;                                                \
    return true;                                                               \
  }

template<>
bool twopass_kernels<double>(int d, int n, bool exact, KernelInfo *sub,
                             KernelInfo *pair)
{
  KM_TWOPASS_CASE(double, 6, 7, 256, 256)
  KM_TWOPASS_CASE(double, 6, 8, 256, 256)
  return false;
}

template<>
bool twopass_kernels<float>(int d, int n, bool exact, KernelInfo *sub,
                            KernelInfo *pair)
{
  KM_TWOPASS_CASE(float, 6, 7, 256, 256)
  KM_TWOPASS_CASE(float, 6, 8, 256, 256)
  return false;
}

#define KM_TWOPASS_FWD_CASE(T, D, N, NT1, NT2)                                 \
  if (d == D && n == N)                                                        \
  {                                                                            \
    pair->fn = (const void *)&kron_pair_fwd<T, D, N, NT1>;                     \
    pair->nt = NT1;                                                            \
    pair->family = FAM_TWOPASS;                                                \
    sub->fn = (const void *)&kron_sub_fwd<T, D, N, NT2>;                       \
    sub->nt = NT2;                                                             \
    sub->family = FAM_TWOPASS;                                                 \
    return true;                                                               \
  }

template<>
bool twopass_fwd_kernels<double>(int d, int n, KernelInfo *pair, KernelInfo *sub)
{
  KM_TWOPASS_FWD_CASE(double, 6, 7, 256, 256)
  KM_TWOPASS_FWD_CASE(double, 6, 8, 256, 256)
  return false;
}

template<>
bool twopass_fwd_kernels<float>(int d, int n, KernelInfo *pair, KernelInfo *sub)
{
  KM_TWOPASS_FWD_CASE(float, 6, 7, 256, 256)
  KM_TWOPASS_FWD_CASE(float, 6, 8, 256, 256)
  return false;
}

} // namespace km
