// kron_lds instantiations for d = 3 (generated list: every n <= 8 whose
// n^d tensor fits the 150 KiB LDS budget of kron_lds).
#include "km_internal.h"

namespace km
{

template<>
bool lds_kernel_d<double, 3>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(double, 3, 1)
    KM_LDS_CASE(double, 3, 2)
    KM_LDS_CASE(double, 3, 3)
    KM_LDS_CASE(double, 3, 4)
    KM_LDS_CASE(double, 3, 5)
    KM_LDS_CASE(double, 3, 6)
    KM_LDS_CASE(double, 3, 7)
    KM_LDS_CASE(double, 3, 8)
  }
  return false;
}

template<>
bool lds_kernel_d<float, 3>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(float, 3, 1)
    KM_LDS_CASE(float, 3, 2)
    KM_LDS_CASE(float, 3, 3)
    KM_LDS_CASE(float, 3, 4)
    KM_LDS_CASE(float, 3, 5)
    KM_LDS_CASE(float, 3, 6)
    KM_LDS_CASE(float, 3, 7)
    KM_LDS_CASE(float, 3, 8)
  }
  return false;
}

} // namespace km
