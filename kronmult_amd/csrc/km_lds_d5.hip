// kron_lds instantiations for d = 5 (generated list: every n <= 8 whose
// n^d tensor fits the 150 KiB LDS budget of kron_lds; float n = 8 fits but
// would need 64 accumulator registers per thread, so it uses kron_slab).
#include "km_internal.h"

namespace km
{

template<>
bool lds_kernel_d<double, 5>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(double, 5, 1)
    KM_LDS_CASE(double, 5, 2)
    KM_LDS_CASE(double, 5, 3)
    KM_LDS_CASE(double, 5, 4)
    KM_LDS_CASE(double, 5, 5)
    KM_LDS_CASE(double, 5, 6)
    KM_LDS_CASE(double, 5, 7)
  }
  return false;
}

template<>
bool lds_kernel_d<float, 5>(int n, bool exact, bool fwd, KernelInfo *out)
{
  switch (n)
  {
    KM_LDS_CASE(float, 5, 1)
    KM_LDS_CASE(float, 5, 2)
    KM_LDS_CASE(float, 5, 3)
    KM_LDS_CASE(float, 5, 4)
    KM_LDS_CASE(float, 5, 5)
    KM_LDS_CASE(float, 5, 6)
    KM_LDS_CASE(float, 5, 7)
  }
  return false;
}

} // namespace km
