// kron_lds instantiations for d = 2 (generated list: every n <= 8 whose
// n^d tensor fits the 150 KiB LDS budget of kron_lds).
#include "km_internal.h"

namespace km
{

template<>
bool lds_kernel_d<double, 2>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(double, 2, 1)
    KM_LDS_CASE(double, 2, 2)
    KM_LDS_CASE(double, 2, 3)
    KM_LDS_CASE(double, 2, 4)
    KM_LDS_CASE(double, 2, 5)
    KM_LDS_CASE(double, 2, 6)
    KM_LDS_CASE(double, 2, 7)
    KM_LDS_CASE(double, 2, 8)
  }
  return false;
}

template<>
bool lds_kernel_d<float, 2>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(float, 2, 1)
    KM_LDS_CASE(float, 2, 2)
    KM_LDS_CASE(float, 2, 3)
    KM_LDS_CASE(float, 2, 4)
    KM_LDS_CASE(float, 2, 5)
    KM_LDS_CASE(float, 2, 6)
    KM_LDS_CASE(float, 2, 7)
    KM_LDS_CASE(float, 2, 8)
  }
  return false;
}

} // namespace km
