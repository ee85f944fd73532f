// kron_lds instantiations for d = 6 (generated list: every n <= 8 whose
// n^d tensor fits the 150 KiB LDS budget of kron_lds).
#include "km_internal.h"

namespace km
{

template<>
bool lds_kernel_d<double, 6>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(double, 6, 1)
    KM_LDS_CASE(double, 6, 2)
    KM_LDS_CASE(double, 6, 3)
    KM_LDS_CASE(double, 6, 4)
    KM_LDS_CASE(double, 6, 5)
  }
  return false;
}

template<>
bool lds_kernel_d<float, 6>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(float, 6, 1)
    KM_LDS_CASE(float, 6, 2)
    KM_LDS_CASE(float, 6, 3)
    KM_LDS_CASE(float, 6, 4)
    KM_LDS_CASE(float, 6, 5)
  }
  return false;
}

} // namespace km
