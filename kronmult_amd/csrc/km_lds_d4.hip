// kron_lds instantiations for d = 4 (generated list: every n <= 8 whose
// n^d tensor fits the 150 KiB LDS budget of kron_lds).
#include "km_internal.h"

namespace km
{

template<>
bool lds_kernel_d<double, 4>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(double, 4, 1)
    KM_LDS_CASE(double, 4, 2)
    KM_LDS_CASE(double, 4, 3)
    KM_LDS_CASE(double, 4, 4)
    KM_LDS_CASE(double, 4, 5)
    KM_LDS_CASE(double, 4, 6)
    KM_LDS_CASE(double, 4, 7)
    KM_LDS_CASE(double, 4, 8)
  }
  return false;
}

template<>
bool lds_kernel_d<float, 4>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(float, 4, 1)
    KM_LDS_CASE(float, 4, 2)
    KM_LDS_CASE(float, 4, 3)
    KM_LDS_CASE(float, 4, 4)
    KM_LDS_CASE(float, 4, 5)
    KM_LDS_CASE(float, 4, 6)
    KM_LDS_CASE(float, 4, 7)
    KM_LDS_CASE(float, 4, 8)
  }
  return false;
}

} // namespace km
