// Pooled xbatched tables built on the device from index lists: the step
// before kronmultN_xbatched in a sparse-grid operator apply, where every
// entry names one matrix per dimension from a pool, one input vector and one
// output vector.  The A table has the layout kronmultN_xbatched reads,
// Aarray[k + d*b] (kronmult_xbatched.hpp:67-69).  One thread per entry;
// an index outside its pool leaves a null pointer and is counted.
#include "km_internal.h"

namespace km
{
namespace
{

template<typename T>
__global__ __launch_bounds__(256) void build_tables(TableArgs<T> t)
{
  for (long long b = (long long)blockIdx.x * 256 + threadIdx.x; b < t.batch;
       b += (long long)gridDim.x * 256)
  {
    int bad = 0;
    for (int k = 0; k < t.d; k++)
    {
      const int ia = gld(t.aidx + k + (long long)t.d * b);
      const bool ok = ia >= 0 && ia < t.a_count[k];
      bad += !ok;
      gst(t.Aarray + k + (long long)t.d * b,
          ok ? t.pool[k] + (long long)ia * t.a_stride[k] : (const T *)nullptr);
    }
    const int ix = gld(t.xidx + b);
    const int iy = gld(t.yidx + b);
    const bool okx = ix >= 0 && ix < t.x_count;
    const bool oky = iy >= 0 && iy < t.y_count;
    gst(t.pX + b, okx ? t.X + (long long)ix * t.x_stride : (T *)nullptr);
    gst(t.pY + b, oky ? t.Y + (long long)iy * t.y_stride : (T *)nullptr);
    bad += !okx + !oky;
    if (bad && t.bad)
      gatomic_add(t.bad, bad);
  }
}

} // namespace

template<typename T>
hipError_t tables_launch(TableArgs<T> &t, int cus, hipStream_t stream)
{
  long long grid = (t.batch + 255) / 256;
  const long long cap = (long long)(cus > 0 ? cus : 256) * 16;
  if (grid > cap)
    grid = cap;
  if (grid < 1)
    grid = 1;
  void *args[] = {&t};
  return hipLaunchKernel((const void *)&build_tables<T>, dim3((unsigned)grid),
                         dim3(256), args, 0, stream);
}

template hipError_t tables_launch<double>(TableArgs<double> &, int, hipStream_t);
template hipError_t tables_launch<float>(TableArgs<float> &, int, hipStream_t);

} // namespace km
