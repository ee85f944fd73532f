// kron_lds instantiations for d = 1 (generated list: every n <= 8 whose
// n^d tensor fits the 150 KiB LDS budget of kron_lds).
#include "km_internal.h"

namespace km
{

template<>
bool lds_kernel_d<double, 1>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(double, 1, 1)
    KM_LDS_CASE(double, 1, 2)
    KM_LDS_CASE(double, 1, 3)
    KM_LDS_CASE(double, 1, 4)
    KM_LDS_CASE(double, 1, 5)
    KM_LDS_CASE(double, 1, 6)
    KM_LDS_CASE(double, 1, 7)
    KM_LDS_CASE(double, 1, 8)
  }
  return false;
}

template<>
bool lds_kernel_d<float, 1>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(float, 1, 1)
    KM_LDS_CASE(float, 1, 2)
    KM_LDS_CASE(float, 1, 3)
    KM_LDS_CASE(float, 1, 4)
    KM_LDS_CASE(float, 1, 5)
    KM_LDS_CASE(float, 1, 6)
    KM_LDS_CASE(float, 1, 7)
    KM_LDS_CASE(float, 1, 8)
  }
  return false;
}

} // namespace km
